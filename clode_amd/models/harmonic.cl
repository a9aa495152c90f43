// Harmonic oscillator x' = omega y, y' = -omega x: from (A, 0),
// x = A cos(omega t), y = -A sin(omega t); period 2 pi/omega.
// Variables (x, y); parameter omega.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype omega = p_[0];
    dx_[0] = omega * x_[1];
    dx_[1] = -omega * x_[0];
}
