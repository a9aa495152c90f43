// Ornstein-Uhlenbeck process dx = theta (mu - x) dt + sigma dW.
// Variable x; parameters (theta, mu, sigma); one Wiener variable.
void getRHS(const realtype t, const realtype x_[], const realtype p_[],
            realtype dx_[], realtype aux_[], const realtype w_[])
{
    realtype theta = p_[0], mu = p_[1], sigma = p_[2];
    dx_[0] = theta * (mu - x_[0]) + sigma * w_[0];
}
