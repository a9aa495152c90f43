// Exponential decay x' = -lambda x: x(t) = x0 exp(-lambda t).
// Variable x; parameter lambda.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    dx_[0] = -p_[0] * x_[0];
}
