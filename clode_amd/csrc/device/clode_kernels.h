// Ensemble drivers: one member per lane, state in registers for the whole
// integration, coalesced structure-of-arrays loads/stores at the ends.
//
//   clode_transient           clode/cpp/transient.cl:9-77
//   clode_initialize_observer clode/cpp/initializeObserver.cl:9-84
//   clode_features            clode/cpp/features.cl:10-106
//   clode_trajectory          clode/cpp/trajectory.cl:14-112
//
// Semantics kept from the reference drivers: the time loop runs while
// t <= tspan[1] (t < tspan[1] in the observer warm-up pass) and
// step < max_steps; the stepper's status flag is ignored; the polar-normal
// cache is reset at every launch; the warm-up pass of two-pass observers
// neither advances the member's noise stream nor its step size, so the
// features pass replays the same noise; the trajectory stores the initial
// point and then every nout-th step.  One deliberate change: the trajectory
// never writes sample index max_store (the reference can write one sample past
// the end of its buffers); nStored still counts that last step.
#ifndef CLODE_KERNELS_H
#define CLODE_KERNELS_H

#include "clode_kernel_args.h"

#ifndef CLODE_BLOCK_MAX
#define CLODE_BLOCK_MAX 256
#endif

#if CLODE_PREC == 3
typedef double xstore;   // x0 / xf: hi - c of the compensated pair
typedef double dtstore;  // dt, tf
typedef double trtstore; // stored trajectory time
#else
typedef realtype xstore;
typedef realtype dtstore;
typedef realtype trtstore;
#endif

struct Member {
    realtype p[N_PAR1];
    realtype xi[N_VAR];
    realtype dxi[N_VAR];
    realtype auxi[N_AUX1];
    realtype wi[N_WIENER1];
#if CLODE_PREC == 3
    realtype xc[N_VAR];
    double dt;
#else
    realtype dt;
#endif
    timetype ti;
    RngData rd;
};

CLODE_DEV SolverParamsR make_solver_params(const ClodeArgs &a)
{
    SolverParamsR sp;
    sp.dt = (realtype)a.sp_dt;
    sp.dtmax = (realtype)a.sp_dtmax;
    sp.abstol = (realtype)a.sp_abstol;
    sp.reltol = (realtype)a.sp_reltol;
    sp.max_steps = a.sp_max_steps;
    sp.max_store = a.sp_max_store;
    sp.nout = a.sp_nout;
    return sp;
}

#ifdef CLODE_OBSERVER
CLODE_DEV ObserverParamsR make_observer_params(const ClodeArgs &a)
{
    ObserverParamsR op;
    op.eVarIx = a.op_eVarIx;
    op.fVarIx = a.op_fVarIx;
    op.maxEventCount = a.op_maxEventCount;
    op.maxEventTimestamps = a.op_maxEventTimestamps;
    op.minXamp = (realtype)a.op_minXamp;
    op.minIMI = (realtype)a.op_minIMI;
    op.nHoodRadius = (realtype)a.op_nHoodRadius;
    op.xUpThresh = (realtype)a.op_xUpThresh;
    op.xDownThresh = (realtype)a.op_xDownThresh;
    op.dxUpThresh = (realtype)a.op_dxUpThresh;
    op.dxDownThresh = (realtype)a.op_dxDownThresh;
    op.eps_dx = (realtype)a.op_eps_dx;
    return op;
}
#endif

CLODE_DEV void load_state(Member &m, const ClodeArgs &a, unsigned int i)
{
    const size_t n = a.nPts;
    const xstore *x0 = (const xstore *)a.x0;
#pragma unroll
    for (int j = 0; j < N_VAR; ++j) {
#if CLODE_PREC == 3
        double v = x0[j * n + i];
        m.xi[j] = (realtype)v;
        m.xc[j] = (realtype)((double)m.xi[j] - v);
#else
        m.xi[j] = x0[j * n + i];
#endif
    }
}

CLODE_DEV void load_member(Member &m, const ClodeArgs &a, unsigned int i)
{
    const size_t n = a.nPts;
    m.dt = ((const dtstore *)a.dt)[i];
    const realtype *pars = (const realtype *)a.pars;
#pragma unroll
    for (int j = 0; j < N_PAR; ++j)
        m.p[j] = pars[j * n + i];
    if (N_PAR == 0)
        m.p[0] = ZERO;
    load_state(m, a, i);
    m.rd.s0 = a.rng[i];
    m.rd.s1 = a.rng[n + i];
    m.rd.useLast = false;
    m.rd.last = ZERO;
#pragma unroll
    for (int j = 0; j < N_AUX1; ++j)
        m.auxi[j] = ZERO;
}

CLODE_DEV void initial_noise(Member &m)
{
    m.wi[0] = ZERO;
#pragma unroll
    for (int j = 0; j < N_WIENER; ++j) {
#if CLODE_STOCHASTIC
#if CLODE_PREC == 3
        m.wi[j] = rng_normal(m.rd) / sqrtf((realtype)m.dt);
#else
        m.wi[j] = rng_normal(m.rd) / sqrt(m.dt);
#endif
#else
        m.wi[j] = ZERO;
#endif
    }
}

CLODE_DEV void store_member(Member &m, const ClodeArgs &a, unsigned int i)
{
    const size_t n = a.nPts;
    xstore *xf = (xstore *)a.xf;
#pragma unroll
    for (int j = 0; j < N_VAR; ++j) {
#if CLODE_PREC == 3
        xf[j * n + i] = (double)m.xi[j] - (double)m.xc[j];
#else
        xf[j * n + i] = m.xi[j];
#endif
    }
    a.rng[i] = m.rd.s0;
    a.rng[n + i] = m.rd.s1;
    ((dtstore *)a.dt)[i] = m.dt;
    ((dtstore *)a.tf)[i] = m.ti;
}

// optional per-member accumulation of member-steps (benchmark accounting)
CLODE_DEV void count_steps(const ClodeArgs &a, unsigned int i, unsigned int step)
{
    if (a.steps)
        a.steps[i] += step;
}

#if CLODE_PREC == 3
#define CLODE_STEP(m)                                                                              \
    stepper(m.ti, m.xi, m.xc, m.dxi, m.p, sp, m.dt, t1, m.auxi, m.wi, m.rd, a.sp_reltol,            \
            a.sp_abstol / a.sp_reltol, a.sp_dtmax)
#else
#define CLODE_STEP(m) stepper(m.ti, m.xi, m.dxi, m.p, sp, m.dt, t1, m.auxi, m.wi, m.rd)
#endif

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) clode_transient(const ClodeArgs a)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nPts)
        return;
    const SolverParamsR sp = make_solver_params(a);
    const timetype t1 = (timetype)a.t1;
    Member m;
    m.ti = (timetype)a.t0;
    load_member(m, a, i);
    initial_noise(m);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
    unsigned int step = 0;
    while (m.ti <= t1 && step < sp.max_steps) {
        ++step;
        CLODE_STEP(m);
    }
    store_member(m, a, i);
    count_steps(a, i, step);
}

#ifdef CLODE_OBSERVER
extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) clode_initialize_observer(const ClodeArgs a)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nPts)
        return;
    const SolverParamsR sp = make_solver_params(a);
    const ObserverParamsR op = make_observer_params(a);
    const timetype t0 = (timetype)a.t0;
    const timetype t1 = (timetype)a.t1;
    ObsIO io{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0};
    Member m;
    m.ti = t0;
    load_member(m, a, i);
    initial_noise(m);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
    Observer ob;
    ob.init(m.ti, m.xi, m.dxi, m.auxi, op, io);
#if CLODE_TWO_PASS
    unsigned int step = 0;
    while (m.ti < t1 && step < sp.max_steps) {
        ++step;
        CLODE_STEP(m);
        ob.warmup(m.ti, m.xi, m.dxi, m.auxi, op);
    }
    m.ti = t0;
    load_state(m, a, i);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
    count_steps(a, i, step);
#endif
    ob.initEventDetector(m.ti, m.xi, m.dxi, m.auxi, op);
    ObsStore st{io};
    ob.fields(st);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) clode_features(const ClodeArgs a)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nPts)
        return;
    const SolverParamsR sp = make_solver_params(a);
    const ObserverParamsR op = make_observer_params(a);
    const timetype t0 = (timetype)a.t0;
    const timetype t1 = (timetype)a.t1;
    ObsIO io{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0};
    Member m;
    m.ti = t0;
    load_member(m, a, i);
    Observer ob;
    {
        ObsLoad ld{io};
        ob.fields(ld);
    }
    initial_noise(m);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
    unsigned int step = 0;
    while (m.ti <= t1 && step < sp.max_steps) {
        ++step;
        CLODE_STEP(m);
        ob.update(m.ti, m.xi, m.dxi, m.auxi, op, io);
        if (ob.eventFunction(m.ti, m.xi, m.dxi, m.auxi, op)) {
            if (ob.computeEventFeatures(m.ti, m.xi, m.dxi, m.auxi, op, io))
                break;
        }
    }
    FeatOut F{(ostore *)a.F, a.nPts, i, 0};
    ob.finalizeFeatures(m.ti, F, io);
    ob.finalizeObserverData(m.ti, t0);
    ObsStore st{io};
    ob.fields(st);
    store_member(m, a, i);
    count_steps(a, i, step);
}
#endif  // CLODE_OBSERVER

CLODE_DEV void trajectory_store(const ClodeArgs &a, unsigned int i, unsigned int s, const Member &m)
{
    const size_t n = a.nPts;
    ((trtstore *)a.tr_t)[(size_t)s * n + i] = (trtstore)m.ti;
    realtype *X = (realtype *)a.tr_x;
    realtype *DX = (realtype *)a.tr_dx;
#pragma unroll
    for (int j = 0; j < N_VAR; ++j) {
        X[((size_t)s * N_VAR + j) * n + i] = m.xi[j];
        DX[((size_t)s * N_VAR + j) * n + i] = m.dxi[j];
    }
    realtype *AUX = (realtype *)a.tr_aux;
#pragma unroll
    for (int j = 0; j < N_AUX; ++j)
        AUX[((size_t)s * N_AUX + j) * n + i] = m.auxi[j];
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) clode_trajectory(const ClodeArgs a)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nPts)
        return;
    const SolverParamsR sp = make_solver_params(a);
    const timetype t1 = (timetype)a.t1;
    Member m;
    m.ti = (timetype)a.t0;
    load_member(m, a, i);
    initial_noise(m);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
    unsigned int storeix = 0;
    trajectory_store(a, i, storeix, m);
    unsigned int step = 0;
    while (m.ti <= t1 && step < sp.max_steps && storeix < sp.max_store) {
        ++step;
        CLODE_STEP(m);
        if (step % sp.nout == 0) {
            ++storeix;
            if (storeix < sp.max_store)
                trajectory_store(a, i, storeix, m);
        }
    }
    a.nStored[i] = (int)storeix;
    store_member(m, a, i);
    count_steps(a, i, step);
}

#endif  // CLODE_KERNELS_H
