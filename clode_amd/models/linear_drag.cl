// Falling body with linear drag: y1' = y2, y2' = (w - k y2)/m.
// From y(0) = 0: y2 = (w/k)(1 - exp(-k t/m)),
//                y1 = (w/k)(t - (m/k)(1 - exp(-k t/m))).
// Parameters (m, w, k, H); aux: y1 - H.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype m = p_[0], w = p_[1], k = p_[2], H = p_[3];
    dx_[0] = x_[1];
    dx_[1] = (w - k * x_[1]) / m;
    aux_[0] = x_[0] - H;
}
