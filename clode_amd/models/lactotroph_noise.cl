// Lactotroph model (see lactotroph.cl) with additive current noise
// sigma * w on the voltage equation.  Parameters: gcal, gsk, gbk, sigma;
// gk fixed at 2.0 nS.  One Wiener variable; aux: ica.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    const realtype Cm = RCONST(10.0), gl = RCONST(0.050), gk = RCONST(2.0);
    const realtype vca = RCONST(60.0), vk = RCONST(-75.0), vl = RCONST(-50.0);
    const realtype vm = RCONST(-20.0), sm = RCONST(12.0);
    const realtype vn = RCONST(-5.0), sn = RCONST(10.0);
    const realtype vf = RCONST(-20.0), sf = RCONST(2.0);
    const realtype taun = RCONST(30.0), taubk = RCONST(8.0);
    const realtype ff = RCONST(0.010), alpha = RCONST(0.00150), kc = RCONST(0.20), ks = RCONST(0.40);
    realtype v = x_[0], n = x_[1], f = x_[2], c = x_[3];
    realtype gcal = p_[0], gsk = p_[1], gbk = p_[2], sigma = p_[3];
    realtype c2 = c * c;
    realtype vkdrive = v - vk;
    realtype minf = RCONST(1.0) / (RCONST(1.0) + exp((vm - v) / sm));
    realtype ninf = RCONST(1.0) / (RCONST(1.0) + exp((vn - v) / sn));
    realtype finf = RCONST(1.0) / (RCONST(1.0) + exp((vf - v) / sf));
    realtype ica = gcal * minf * (v - vca);
    realtype isk = gsk * c2 / (c2 + ks * ks) * vkdrive;
    realtype ibk = gbk * f * vkdrive;
    realtype ik = gk * n * vkdrive;
    realtype il = gl * (v - vl);
    realtype inoise = sigma * w_[0];
    realtype itot = ica + isk + ibk + ik + il + inoise;
    dx_[0] = -itot / Cm;
    dx_[1] = (ninf - n) / taun;
    dx_[2] = (finf - f) / taubk;
    dx_[3] = -ff * (alpha * ica + kc * c);
    aux_[0] = ica;
}
