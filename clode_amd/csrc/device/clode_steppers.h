// Time steppers, one ensemble member per lane, all state in registers.
//
// Registry (name -> CLODE_STEPPER id), as clode/cpp/steppers.cl:26-43:
//   euler=1  heun=2  rk4=3  bs23=4  dopri5=5  seuler=6 (Euler-Maruyama)
//
// Plain modes (CLODE_PREC 1 = single, 2 = double) restate the reference's
// arithmetic expression by expression so that, compiled with
// -ffp-contract=on, every multiply-add is fused exactly where the reference's
// OpenCL C would be:
//   fixed-step wrapper  clode/cpp/steppers/fixed_explicit_step.clh:9-36
//   Euler               clode/cpp/steppers/fixed_explicit_Euler.clh:5-17
//   Heun                clode/cpp/steppers/fixed_explicit_Trapezoidal.clh:5-22
//   RK4                 clode/cpp/steppers/fixed_explicit_RK4.clh:5-39
//   adaptive wrapper    clode/cpp/steppers/adaptive_explicit_step.clh:4-80
//   Bogacki-Shampine    clode/cpp/steppers/adaptive_bs23.clh:3-62
//   Dormand-Prince      clode/cpp/steppers/adaptive_dp45.clh:4-110
//   seuler = Euler + fresh Wiener increments after each update
//                       (clode/cpp/steppers.cl:54-57)
// Tableau constants are written as quotients of realtype literals so that
// single precision folds them to the correctly rounded float quotient.
//
// fp32-compensated mode (CLODE_PREC 3) keeps the same tableaux but carries
// time and the step size in double, adds each state increment with Kahan
// compensation (x, c) and drives the controller from a double error norm
// and a double accept/reject test; stage algebra, the right-hand side and
// the embedded error combination h sum(E_i k_i) run in float.  That
// combination rounds to about one float ulp of h sum(|E_i k_i|), the same
// size as the rounding the float stage derivatives k_i already carry, so
// forming it in double would not sharpen the estimate.
#ifndef CLODE_STEPPERS_H
#define CLODE_STEPPERS_H

#if CLODE_STEPPER == 1 || CLODE_STEPPER == 2 || CLODE_STEPPER == 3 || CLODE_STEPPER == 6
#define CLODE_FIXED_STEP 1
#elif CLODE_STEPPER == 4 || CLODE_STEPPER == 5 || CLODE_STEPPER == 7
#define CLODE_ADAPTIVE_STEP 1
#else
#error "unknown CLODE_STEPPER"
#endif

#if CLODE_STEPPER == 6
#define CLODE_STOCHASTIC 1
#endif

// End of the adaptive retry loop.  The reference's loop
// (adaptive_explicit_step.clh:20-62) ends when a rejected trial's step is at
// most hmin.  After a rejection the next trial's (clamped) step is strictly
// smaller until it is clamped at hmin, so a trial equal to the previous one
// can only happen at hmin, when t + hmin rounds to a step longer than hmin
// (t just below a power of two): that trial would repeat, and be rejected,
// forever.  It is failed at hmin instead (stall detection).  Every sequence
// that ends in the reference ends here after the same trials.  As a second
// guard a step makes at most CLODE_MAX_ATTEMPTS trials; no sequence that
// ends in the reference gets near it (from the largest double down to hmin
// takes about 2300 trials, in float about 300).
#define CLODE_MAX_ATTEMPTS 4096

#define N_PAR1 (N_PAR > 0 ? N_PAR : 1)
#define N_AUX1 (N_AUX > 0 ? N_AUX : 1)
#define N_WIENER1 (N_WIENER > 0 ? N_WIENER : 1)

// Solver parameters in working precision (clode/cpp/clODE_struct_defs.cl:11-20)
struct SolverParamsR {
    realtype dt;
    realtype dtmax;
    realtype abstol;
    realtype reltol;
    unsigned int max_steps;
    unsigned int max_store;
    unsigned int nout;
};

// The drivers call step_attempt() once per loop iteration.  A fixed-step
// stepper completes its step in one attempt.  The adaptive controller makes
// one trial per attempt and keeps what its retry loop carries between trials
// (adaptive_explicit_step.clh:14-62: the trial step size, the attempt count
// and whether a trial has failed) in StepCtl, so a lane whose trial was
// rejected retries on the next iteration while the other lanes of its wave go
// on with their next steps.  Each member still sees exactly the reference's
// sequence of trials; only the interleaving across lanes changes.
#if CLODE_PREC == 3
typedef double ctlreal;
#else
typedef realtype ctlreal;
#endif
struct StepCtl {
    ctlreal newDt;   // the next trial's step before clamping
    ctlreal lastDt;  // the previous trial's clamped step (stall detection)
    int attempts;
    bool noFailedSteps;
    bool inStep;
};
CLODE_DEV void step_ctl_reset(StepCtl &c)
{
    c.newDt = 0;
    c.lastDt = 0;
    c.attempts = 0;
    c.noFailedSteps = true;
    c.inStep = false;
}

#if CLODE_PREC != 3
// =====================================================================
//  Plain single / double precision (reference semantics)
// =====================================================================

#if CLODE_STEPPER == 1 || CLODE_STEPPER == 6
CLODE_DEV void do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                       const realtype dt, realtype aux[], const realtype wi[])
{
    for (int k = 0; k < N_VAR; k++)
        xi[k] += dt * k1[k];
    ti += dt;
}
#endif

#if CLODE_STEPPER == 2
CLODE_DEV void do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                       const realtype dt, realtype aux[], const realtype wi[])
{
    realtype th = ti + dt;
    realtype tmp[N_VAR], k2[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = fma(dt, k1[k], xi[k]);
    getRHS(th, tmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xi[k] += dt * RCONST(0.5) * (k1[k] + k2[k]);
    ti = th;
}
#endif

#if CLODE_STEPPER == 3
CLODE_DEV void do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                       const realtype dt, realtype aux[], const realtype wi[])
{
    realtype tmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR];
    realtype h2 = dt * RCONST(0.5);
    realtype th2 = ti + h2;
    realtype th = ti + dt;
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + h2 * k1[k];
    getRHS(th2, tmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + h2 * k2[k];
    getRHS(th2, tmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + dt * k3[k];
    getRHS(th, tmp, pars, k4, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xi[k] += dt * (k1[k] + RCONST(2.0) * k2[k] + RCONST(2.0) * k3[k] + k4[k]) / RCONST(6.0);
    ti = th;
}
#endif

#if CLODE_FIXED_STEP
// Fixed-step wrapper: dt never changes; k1 is re-evaluated at the new point
// (after a fresh Wiener draw for seuler) and reused as the next step's k1.
CLODE_DEV bool step_attempt(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                            const SolverParamsR &sp, realtype &dt, const realtype t1,
                            realtype aux[], realtype wi[], RngData &rd, StepCtl &)
{
    do_step(ti, xi, k1, pars, dt, aux, wi);
#if CLODE_STOCHASTIC
    for (int j = 0; j < N_WIENER; ++j)
        wi[j] = rng_normal(rd) / sqrt(dt);
#endif
    getRHS(ti, xi, pars, k1, aux, wi);
    return true;
}
#endif

#if CLODE_STEPPER == 7
// Heun-Euler 1(2) (adaptive_he12.clh:9-40, selected by HEUN_EULER in
// steppers.cl:89-91 but not in the reference's name map): an Euler
// predictor, the trapezoidal average as the new state, half the slope
// difference as the error; the slope at the new state is evaluated again.
#define CLODE_ROOT_N 2  // 1 / (error order 1 + 1): the square root
#define CLODE_MAX_SHRINK RCONST(0.5)
#define CLODE_MAX_GROW RCONST(5.0)
CLODE_DEV realtype do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                           const realtype dt, realtype aux[], realtype err[], const realtype wi[])
{
    realtype tNew = ti + dt;
    realtype newDt = tNew - ti;
    realtype xtmp[N_VAR], k2[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * k1[k];
    getRHS(tNew, xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xi[k] = xi[k] + newDt * RCONST(0.5) * (k1[k] + k2[k]);
    for (int k = 0; k < N_VAR; k++)
        err[k] = newDt * RCONST(0.5) * (-k1[k] + k2[k]);
    getRHS(tNew, xi, pars, k1, aux, wi);
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_STEPPER == 4
#define CLODE_ROOT_N 3  // 1 / (error order 2 + 1)
#define CLODE_MAX_SHRINK RCONST(0.5)
#define CLODE_MAX_GROW RCONST(5.0)
CLODE_DEV realtype do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                           const realtype dt, realtype aux[], realtype err[], const realtype wi[])
{
    const realtype B1 = RCONST(2.0) / RCONST(9.0);
    const realtype B2 = RCONST(1.0) / RCONST(3.0);
    const realtype B3 = RCONST(4.0) / RCONST(9.0);
    const realtype E1 = RCONST(-5.0) / RCONST(72.0);
    const realtype E2 = RCONST(1.0) / RCONST(12.0);
    const realtype E3 = RCONST(1.0) / RCONST(9.0);
    const realtype E4 = RCONST(-1.0) / RCONST(8.0);
    realtype tNew = ti + dt;
    realtype newDt = tNew - ti;
    realtype xtmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * RCONST(0.5) * k1[k];
    getRHS(ti + newDt * RCONST(0.5), xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * RCONST(0.75) * k2[k];
    getRHS(ti + newDt * RCONST(0.75), xtmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xi[k] = xi[k] + newDt * (B1 * k1[k] + B2 * k2[k] + B3 * k3[k]);
    getRHS(tNew, xi, pars, k4, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        err[k] = newDt * (E1 * k1[k] + E2 * k2[k] + E3 * k3[k] + E4 * k4[k]);
        k1[k] = k4[k];
    }
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_STEPPER == 5
#define CLODE_ROOT_N 5  // 1 / (error order 4 + 1)
#define CLODE_MAX_SHRINK RCONST(0.1)
#define CLODE_MAX_GROW RCONST(5.0)
CLODE_DEV realtype do_step(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                           const realtype dt, realtype aux[], realtype err[], const realtype wi[])
{
    const realtype A2 = RCONST(1.0) / RCONST(5.0);
    const realtype A3 = RCONST(3.0) / RCONST(10.0);
    const realtype A4 = RCONST(4.0) / RCONST(5.0);
    const realtype A5 = RCONST(8.0) / RCONST(9.0);
    const realtype B21 = RCONST(1.0) / RCONST(5.0);
    const realtype B31 = RCONST(3.0) / RCONST(40.0);
    const realtype B32 = RCONST(9.0) / RCONST(40.0);
    const realtype B41 = RCONST(44.0) / RCONST(45.0);
    const realtype B42 = RCONST(-56.0) / RCONST(15.0);
    const realtype B43 = RCONST(32.0) / RCONST(9.0);
    const realtype B51 = RCONST(19372.0) / RCONST(6561.0);
    const realtype B52 = RCONST(-25360.0) / RCONST(2187.0);
    const realtype B53 = RCONST(64448.0) / RCONST(6561.0);
    const realtype B54 = RCONST(-212.0) / RCONST(729.0);
    const realtype B61 = RCONST(9017.0) / RCONST(3168.0);
    const realtype B62 = RCONST(-355.0) / RCONST(33.0);
    const realtype B63 = RCONST(46732.0) / RCONST(5247.0);
    const realtype B64 = RCONST(49.0) / RCONST(176.0);
    const realtype B65 = RCONST(-5103.0) / RCONST(18656.0);
    const realtype C1 = RCONST(35.0) / RCONST(384.0);
    const realtype C3 = RCONST(500.0) / RCONST(1113.0);
    const realtype C4 = RCONST(125.0) / RCONST(192.0);
    const realtype C5 = RCONST(-2187.0) / RCONST(6784.0);
    const realtype C6 = RCONST(11.0) / RCONST(84.0);
    const realtype E1 = RCONST(71.0) / RCONST(57600.0);
    const realtype E3 = RCONST(-71.0) / RCONST(16695.0);
    const realtype E4 = RCONST(71.0) / RCONST(1920.0);
    const realtype E5 = RCONST(-17253.0) / RCONST(339200.0);
    const realtype E6 = RCONST(22.0) / RCONST(525.0);
    const realtype E7 = RCONST(-1.0) / RCONST(40.0);
    realtype tNew = ti + dt;
    realtype newDt = tNew - ti;
    realtype xtmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR], k5[N_VAR], k6[N_VAR], k7[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * (B21 * k1[k]);
    getRHS(ti + A2 * newDt, xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * (B31 * k1[k] + B32 * k2[k]);
    getRHS(ti + A3 * newDt, xtmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * (B41 * k1[k] + B42 * k2[k] + B43 * k3[k]);
    getRHS(ti + A4 * newDt, xtmp, pars, k4, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * (B51 * k1[k] + B52 * k2[k] + B53 * k3[k] + B54 * k4[k]);
    getRHS(ti + A5 * newDt, xtmp, pars, k5, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + newDt * (B61 * k1[k] + B62 * k2[k] + B63 * k3[k] + B64 * k4[k] + B65 * k5[k]);
    getRHS(ti + newDt, xtmp, pars, k6, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xi[k] = xi[k] + newDt * (C1 * k1[k] + C3 * k3[k] + C4 * k4[k] + C5 * k5[k] + C6 * k6[k]);
    getRHS(tNew, xi, pars, k7, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        err[k] = newDt * (E1 * k1[k] + E3 * k3[k] + E4 * k4[k] + E5 * k5[k] + E6 * k6[k] + E7 * k7[k]);
        k1[k] = k7[k];
    }
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_ADAPTIVE_STEP
// Error-per-unit-step controller (MATLAB ode23/ode45 style): elementwise
// relative error against max(|x|, |xnew|, abstol/reltol), infinity norm,
// first rejection shrinks by max(MAX_SHRINK, 0.8 (rtol/err)^(1/(p+1))) (the
// root is cm::root<p+1>, within 2 ulp of the reference's pow),
// later rejections halve, growth after an unrejected step is capped at
// MAX_GROW, and the step is clipped to the end of the span and to
// [hmin, dtmax] with hmin = 16 ulp(t).  One call makes one trial: it returns
// false when the trial was rejected (state unchanged, to be retried), true
// when the step is done, either accepted or failed at hmin (state unchanged,
// dt = hmin; the drivers do not distinguish the two, as the reference's do
// not look at its -1 flag).
CLODE_DEV bool step_attempt(realtype &ti, realtype xi[], realtype k1[], const realtype pars[],
                            const SolverParamsR &sp, realtype &dt, const realtype t1,
                            realtype aux[], realtype wi[], RngData &rd, StepCtl &c)
{
    realtype tNew, normErr, err[N_VAR], newxi[N_VAR], newk1[N_VAR];
    realtype threshold = sp.abstol / sp.reltol;
    realtype hmin = cm::min_step(ti, RCONST(1.1) * t1);  // 16 |fabs(nextafter(ti, 1.1 t1)) - ti|
    if (!c.inStep) {
        c.inStep = true;
        c.noFailedSteps = true;
        c.attempts = 0;
        c.newDt = dt;
    }
    realtype newDt = fmin(fmax(c.newDt, hmin), sp.dtmax);
    if ((!c.noFailedSteps && newDt == c.lastDt) || ++c.attempts > CLODE_MAX_ATTEMPTS) {
        dt = hmin;  // stalled at hmin: the reference would retry this trial forever
        c.inStep = false;
        return true;
    }
    c.lastDt = newDt;
    tNew = ti;
    for (int j = 0; j < N_VAR; j++) {
        newxi[j] = xi[j];
        newk1[j] = k1[j];
    }
    newDt = do_step(tNew, newxi, newk1, pars, newDt, aux, err, wi);
    for (int j = 0; j < N_VAR; j++)
        err[j] /= fmax(fmax(fabs(xi[j]), fabs(newxi[j])), threshold);
    normErr = ZERO;
    for (int j = 0; j < N_VAR; j++)
        normErr = fmax(fabs(err[j]), normErr);
    // the factor both outcomes use, formed once for the whole wave:
    // 0.8 (rtol/err)^(1/(p+1)) (the reference's pow(rtol/err, 1/(p+1)),
    // here the division-free fixed root shared with the CPU oracle)
    const realtype factor = RCONST(0.8) * cm::root<CLODE_ROOT_N>(sp.reltol / normErr);
    if (normErr > sp.reltol) {
        if (newDt <= hmin) {
            dt = hmin;
            c.inStep = false;
            return true;
        }
        newDt *= c.noFailedSteps ? fmax(CLODE_MAX_SHRINK, factor) : RCONST(0.5);
        c.noFailedSteps = false;
        c.newDt = newDt;
        return false;
    }
    if (c.noFailedSteps)
        newDt *= fmin(CLODE_MAX_GROW, factor);
    newDt = fmin(newDt, t1 - tNew);
    newDt = fmin(fmax(newDt, hmin), sp.dtmax);
    dt = newDt;
    ti = tNew;
    for (int j = 0; j < N_VAR; j++) {
        xi[j] = newxi[j];
        k1[j] = newk1[j];
    }
    c.inStep = false;
    return true;
}
#endif

#else  // CLODE_PREC == 3
// =====================================================================
//  fp32-compensated: float state + Kahan compensation, double time/dt,
//  float embedded-error estimate, double error norm
// =====================================================================

// x + d with Kahan compensation c (c holds the negated lost low part)
CLODE_DEV void kahan_add(realtype &x, realtype &c, realtype d)
{
    realtype y = d - c;
    realtype t = x + y;
    c = (t - x) - y;
    x = t;
}

#if CLODE_STEPPER == 1 || CLODE_STEPPER == 6
CLODE_DEV void do_step(double &ti, realtype xi[], realtype xc[], realtype k1[], const realtype pars[],
                       const double dt, realtype aux[], const realtype wi[])
{
    const realtype h = (realtype)dt;
    for (int k = 0; k < N_VAR; k++)
        kahan_add(xi[k], xc[k], h * k1[k]);
    ti += dt;
}
#endif

#if CLODE_STEPPER == 2
CLODE_DEV void do_step(double &ti, realtype xi[], realtype xc[], realtype k1[], const realtype pars[],
                       const double dt, realtype aux[], const realtype wi[])
{
    const realtype h = (realtype)dt;
    double th = ti + dt;
    realtype tmp[N_VAR], k2[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = fmaf(h, k1[k], xi[k]);
    getRHS((realtype)th, tmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        kahan_add(xi[k], xc[k], h * 0.5f * (k1[k] + k2[k]));
    ti = th;
}
#endif

#if CLODE_STEPPER == 3
CLODE_DEV void do_step(double &ti, realtype xi[], realtype xc[], realtype k1[], const realtype pars[],
                       const double dt, realtype aux[], const realtype wi[])
{
    const realtype h = (realtype)dt;
    realtype tmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR];
    realtype h2 = h * 0.5f;
    realtype th2 = (realtype)(ti + 0.5 * dt);
    double th = ti + dt;
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + h2 * k1[k];
    getRHS(th2, tmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + h2 * k2[k];
    getRHS(th2, tmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        tmp[k] = xi[k] + h * k3[k];
    getRHS((realtype)th, tmp, pars, k4, aux, wi);
    const realtype sixth = 1.0f / 6.0f;
    for (int k = 0; k < N_VAR; k++)
        kahan_add(xi[k], xc[k], h * sixth * (k1[k] + 2.0f * k2[k] + 2.0f * k3[k] + k4[k]));
    ti = th;
}
#endif

#if CLODE_FIXED_STEP
CLODE_DEV bool step_attempt(double &ti, realtype xi[], realtype xc[], realtype k1[], const realtype pars[],
                            const SolverParamsR &sp, double &dt, const double t1,
                            realtype aux[], realtype wi[], RngData &rd, const double, const double, const double,
                            StepCtl &)
{
    do_step(ti, xi, xc, k1, pars, dt, aux, wi);
#if CLODE_STOCHASTIC
    const realtype rsq = 1.0f / sqrtf((realtype)dt);
    for (int j = 0; j < N_WIENER; ++j)
        wi[j] = rng_normal(rd) * rsq;
#endif
    getRHS((realtype)ti, xi, pars, k1, aux, wi);
    return true;
}
#endif

#if CLODE_STEPPER == 7
#define CLODE_EXPON 0.5f
#define CLODE_MAX_SHRINK 0.5
#define CLODE_MAX_GROW 5.0
// Heun-Euler 1(2) in float stages; returns the step taken, newxi holds the
// new state without the compensation term, dx its increment, err the error
CLODE_DEV double do_step(double &ti, realtype xi[], realtype k1[], const realtype pars[],
                         const double dt, realtype aux[], realtype err[], realtype dx[], const realtype wi[])
{
    double tNew = ti + dt;
    double newDt = tNew - ti;
    const realtype h = (realtype)newDt;
    realtype xtmp[N_VAR], k2[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * k1[k];
    getRHS((realtype)tNew, xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        dx[k] = h * 0.5f * (k1[k] + k2[k]);
        err[k] = h * 0.5f * (-k1[k] + k2[k]);
        xi[k] = xi[k] + dx[k];
    }
    getRHS((realtype)tNew, xi, pars, k1, aux, wi);
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_STEPPER == 4
#define CLODE_EXPON (1.0f / 3.0f)
#define CLODE_MAX_SHRINK 0.5
#define CLODE_MAX_GROW 5.0
// returns the step taken; newxi holds the 3rd-order solution without the
// compensation term; err the error estimate
CLODE_DEV double do_step(double &ti, realtype xi[], realtype k1[], const realtype pars[],
                         const double dt, realtype aux[], realtype err[], realtype dx[], const realtype wi[])
{
    const realtype B1 = 2.0f / 9.0f, B2 = 1.0f / 3.0f, B3 = 4.0f / 9.0f;
    const realtype E1 = -5.0f / 72.0f, E2 = 1.0f / 12.0f, E3 = 1.0f / 9.0f, E4 = -1.0f / 8.0f;
    double tNew = ti + dt;
    double newDt = tNew - ti;
    const realtype h = (realtype)newDt;
    realtype xtmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * 0.5f * k1[k];
    getRHS((realtype)(ti + 0.5 * newDt), xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * 0.75f * k2[k];
    getRHS((realtype)(ti + 0.75 * newDt), xtmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        dx[k] = h * (B1 * k1[k] + B2 * k2[k] + B3 * k3[k]);
        xtmp[k] = xi[k] + dx[k];
    }
    getRHS((realtype)tNew, xtmp, pars, k4, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        err[k] = h * (E1 * k1[k] + E2 * k2[k] + E3 * k3[k] + E4 * k4[k]);
        k1[k] = k4[k];
        xi[k] = xtmp[k];
    }
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_STEPPER == 5
#define CLODE_EXPON 0.2f
#define CLODE_MAX_SHRINK 0.1
#define CLODE_MAX_GROW 5.0
CLODE_DEV double do_step(double &ti, realtype xi[], realtype k1[], const realtype pars[],
                         const double dt, realtype aux[], realtype err[], realtype dx[], const realtype wi[])
{
    const realtype B21 = 1.0f / 5.0f;
    const realtype B31 = 3.0f / 40.0f, B32 = 9.0f / 40.0f;
    const realtype B41 = 44.0f / 45.0f, B42 = -56.0f / 15.0f, B43 = 32.0f / 9.0f;
    const realtype B51 = 19372.0f / 6561.0f, B52 = -25360.0f / 2187.0f, B53 = 64448.0f / 6561.0f,
                   B54 = -212.0f / 729.0f;
    const realtype B61 = 9017.0f / 3168.0f, B62 = -355.0f / 33.0f, B63 = 46732.0f / 5247.0f,
                   B64 = 49.0f / 176.0f, B65 = -5103.0f / 18656.0f;
    const realtype C1 = 35.0f / 384.0f, C3 = 500.0f / 1113.0f, C4 = 125.0f / 192.0f,
                   C5 = -2187.0f / 6784.0f, C6 = 11.0f / 84.0f;
    const realtype E1 = 71.0f / 57600.0f, E3 = -71.0f / 16695.0f, E4 = 71.0f / 1920.0f,
                   E5 = -17253.0f / 339200.0f, E6 = 22.0f / 525.0f, E7 = -1.0f / 40.0f;
    double tNew = ti + dt;
    double newDt = tNew - ti;
    const realtype h = (realtype)newDt;
    realtype xtmp[N_VAR], k2[N_VAR], k3[N_VAR], k4[N_VAR], k5[N_VAR], k6[N_VAR], k7[N_VAR];
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * (B21 * k1[k]);
    getRHS((realtype)(ti + 0.2 * newDt), xtmp, pars, k2, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * (B31 * k1[k] + B32 * k2[k]);
    getRHS((realtype)(ti + 0.3 * newDt), xtmp, pars, k3, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * (B41 * k1[k] + B42 * k2[k] + B43 * k3[k]);
    getRHS((realtype)(ti + 0.8 * newDt), xtmp, pars, k4, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * (B51 * k1[k] + B52 * k2[k] + B53 * k3[k] + B54 * k4[k]);
    getRHS((realtype)(ti + (8.0 / 9.0) * newDt), xtmp, pars, k5, aux, wi);
    for (int k = 0; k < N_VAR; k++)
        xtmp[k] = xi[k] + h * (B61 * k1[k] + B62 * k2[k] + B63 * k3[k] + B64 * k4[k] + B65 * k5[k]);
    getRHS((realtype)tNew, xtmp, pars, k6, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        dx[k] = h * (C1 * k1[k] + C3 * k3[k] + C4 * k4[k] + C5 * k5[k] + C6 * k6[k]);
        xtmp[k] = xi[k] + dx[k];
    }
    getRHS((realtype)tNew, xtmp, pars, k7, aux, wi);
    for (int k = 0; k < N_VAR; k++) {
        err[k] = h * (E1 * k1[k] + E3 * k3[k] + E4 * k4[k] + E5 * k5[k] + E6 * k6[k] + E7 * k7[k]);
        k1[k] = k7[k];
        xi[k] = xtmp[k];
    }
    ti = tNew;
    return newDt;
}
#endif

#if CLODE_ADAPTIVE_STEP
// Same controller as the plain modes, one trial per call.  The error norm
// max_j |err_j| / max(|x_j|, |x_j'|, atol/rtol) and its test against rtol
// are double, from float operands: |err_j| times the hardware reciprocal of
// the float scale (within 1 ulp; the correctly rounded reciprocal cost four
// instructions more per variable and changed no feature), widened exactly;
// the growth factor (rtol/err)^(1/(p+1)) is formed in float with the hardware
// reciprocal, log2 and exp2 (it only steers the next step size, which stays
// in double).
CLODE_DEV bool step_attempt(double &ti, realtype xi[], realtype xc[], realtype k1[], const realtype pars[],
                            const SolverParamsR &sp, double &dt, const double t1,
                            realtype aux[], realtype wi[], RngData &rd, const double reltol,
                            const realtype threshold, const double dtmax, StepCtl &c)
{
    double tNew, normErr;
    realtype err[N_VAR], newxi[N_VAR], newk1[N_VAR], dx[N_VAR];
    double hmin = cm::min_step(ti, 1.1 * t1);  // 16 |fabs(nextafter(ti, 1.1 t1)) - ti|
    if (!c.inStep) {
        c.inStep = true;
        c.noFailedSteps = true;
        c.attempts = 0;
        c.newDt = dt;
    }
    double newDt = fmin(fmax(c.newDt, hmin), dtmax);
    if ((!c.noFailedSteps && newDt == c.lastDt) || ++c.attempts > CLODE_MAX_ATTEMPTS) {
        dt = hmin;  // stalled at hmin (see CLODE_MAX_ATTEMPTS)
        c.inStep = false;
        return true;
    }
    c.lastDt = newDt;
    tNew = ti;
    for (int j = 0; j < N_VAR; j++) {
        newxi[j] = xi[j];
        newk1[j] = k1[j];
    }
    newDt = do_step(tNew, newxi, newk1, pars, newDt, aux, err, dx, wi);
    normErr = 0.0;
    for (int j = 0; j < N_VAR; j++) {
        realtype sc = fmaxf(fmaxf(fabsf(xi[j]), fabsf(newxi[j])), threshold);
        normErr = fmax((double)fabsf(err[j]) * (double)__builtin_amdgcn_rcpf(sc), normErr);
    }
    // the factor both outcomes use (a rejected trial has normErr > reltol > 0);
    // q is clamped to normal floats so that the hardware log2 sees no
    // denormal, zero or infinity
    const realtype q =
        fminf(fmaxf((realtype)reltol * __builtin_amdgcn_rcpf((realtype)normErr), 1.0e-30f), 1.0e30f);
    const double factor = (double)(0.8f * __builtin_amdgcn_exp2f(CLODE_EXPON * __builtin_amdgcn_logf(q)));
    if (normErr > reltol) {
        if (newDt <= hmin) {
            dt = hmin;
            c.inStep = false;
            return true;
        }
        newDt *= c.noFailedSteps ? fmax(CLODE_MAX_SHRINK, factor) : 0.5;
        c.noFailedSteps = false;
        c.newDt = newDt;
        return false;
    }
    if (c.noFailedSteps)
        newDt *= fmin(CLODE_MAX_GROW, factor);
    newDt = fmin(newDt, t1 - tNew);
    newDt = fmin(fmax(newDt, hmin), dtmax);
    dt = newDt;
    ti = tNew;
    for (int j = 0; j < N_VAR; j++) {
        // newxi = xi + dx (rounded); redo the addition with compensation
        kahan_add(xi[j], xc[j], dx[j]);
        k1[j] = newk1[j];
    }
    c.inStep = false;
    return true;
}
#endif

#endif  // CLODE_PREC

#endif  // CLODE_STEPPERS_H
