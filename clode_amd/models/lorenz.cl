// Lorenz system: x' = sigma (y - x), y' = rho x - y - x z, z' = x y - beta z.
// Variables (x, y, z); parameters (sigma, rho, beta); no aux.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype x = x_[0];
    realtype y = x_[1];
    realtype z = x_[2];
    realtype sigma = p_[0];
    realtype rho = p_[1];
    realtype beta = p_[2];
    dx_[0] = sigma * (y - x);
    dx_[1] = rho * x - y - x * z;
    dx_[2] = x * y - beta * z;
}
