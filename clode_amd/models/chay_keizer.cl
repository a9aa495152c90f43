// Chay-Keizer pancreatic beta-cell burster: voltage v (mV), K+ activation n,
// cytosolic calcium c (uM).  Parameters: gca, gkca, kpmca.  Calcium-activated
// K+ current with Hill-type activation c^2/(c^2 + kd^2).
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    const realtype gk = RCONST(3000.0), vca = RCONST(25.0), vk = RCONST(-75.0);
    const realtype cm = RCONST(5300.0), alpha = RCONST(4.5e-6), fcyt = RCONST(0.01);
    const realtype kd = RCONST(0.4), vm = RCONST(-20.0), sm = RCONST(12.0);
    const realtype vn = RCONST(-16.0), sn = RCONST(5.0), taun = RCONST(20.0);
    realtype v = x_[0], n = x_[1], c = x_[2];
    realtype gca = p_[0], gkca = p_[1], kpmca = p_[2];
    realtype minf = RCONST(1.0) / (RCONST(1.0) + exp((vm - v) / sm));
    realtype ninf = RCONST(1.0) / (RCONST(1.0) + exp((vn - v) / sn));
    realtype omega = pown(c, 2) / (pown(c, 2) + pown(kd, 2));
    realtype ica = gca * minf * (v - vca);
    realtype ik = gk * n * (v - vk);
    realtype ikca = gkca * omega * (v - vk);
    dx_[0] = -(ica + ik + ikca) / cm;
    dx_[1] = (ninf - n) / taun;
    dx_[2] = fcyt * (-alpha * ica - kpmca * c);
}
