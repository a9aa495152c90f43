// x' = cos(omega t): x(t) = sin(omega t)/omega from x(0) = 0.
// Variable x; parameter omega.  Aux: x + 1, the constants 1 and -2.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype omega = p_[0];
    dx_[0] = cos(t * omega);
    aux_[0] = x_[0] + RCONST(1.0);
    aux_[1] = RCONST(1.0);
    aux_[2] = RCONST(-2.0);
}
