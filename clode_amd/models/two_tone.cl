// Two harmonic oscillators, of frequency 1 and 2, and their weighted sum as a
// fifth state: from (1, 0, 1, 0, 1 + b), u1 = cos t, v1 = -sin t,
// u2 = cos 2t, v2 = -sin 2t and s = cos t + b cos 2t (period 2 pi).  For
// b > 1/4, s has two local maxima per period (1 + b at t = 0 and b - 1 at
// t = pi) and two minima (at cos t = -1/(4b)).
// Variables (u1, v1, u2, v2, s); parameter b.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype b = p_[0];
    dx_[0] = x_[1];
    dx_[1] = -x_[0];
    dx_[2] = RCONST(2.0) * x_[3];
    dx_[3] = RCONST(-2.0) * x_[2];
    dx_[4] = x_[1] + RCONST(2.0) * b * x_[3];
}
