// Ensemble drivers: one member per lane, state in registers for the whole
// integration, coalesced structure-of-arrays loads/stores at the ends.
//
//   TransientLane  clode/cpp/transient.cl:9-77
//   InitObsLane    clode/cpp/initializeObserver.cl:9-84
//   FeaturesLane   clode/cpp/features.cl:10-106
//   TrajectoryLane clode/cpp/trajectory.cl:14-112
//
// Each driver is a "lane" with begin (load member i), running (the loop
// condition), advance (one loop iteration: one step attempt, plus the
// driver's per-step work when the attempt completes the step) and finish
// (store member i).  A rejected adaptive trial leaves the loop condition
// unchanged, so the lane simply retries on its next iteration.  The
// grid kernels run one member per lane to completion; the work-queue kernels
// (run_queue) refill a lane with the next unclaimed member as soon as its
// member finishes, so members of very different lengths keep every lane busy.
// Both run the same lane code, so their results are identical.
//
// Semantics kept from the reference drivers: the time loop runs while
// t <= tspan[1] (t < tspan[1] in the observer warm-up pass) and
// step < max_steps; the stepper's status flag is ignored; the polar-normal
// cache is reset at every launch; the warm-up pass of two-pass observers
// neither advances the member's noise stream nor its step size, so the
// features pass replays the same noise; the trajectory stores the initial
// point and then every nout-th step.  One deliberate change: the trajectory
// never writes sample index max_store (the reference can write one sample past
// the end of its buffers); nStored still counts that last step.  Optional
// per-member time spans (tspan_m, [2][nPts]) replace the global span member by
// member; the reference has one span per launch.
//
// The grid kernels are not emitted when CLODE_BATCH is defined: a batched
// program compiles these drivers once per system, each inside its own
// namespace, and generates one dispatching work-queue kernel.
#ifndef CLODE_KERNELS_H
#define CLODE_KERNELS_H

#include "clode_kernel_args.h"

#ifndef CLODE_BLOCK_MAX
#define CLODE_BLOCK_MAX 256
#endif
// extra kernel attributes (code-generation tuning), empty by default
#ifndef CLODE_KERNEL_ATTRS
#define CLODE_KERNEL_ATTRS
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
    StepCtl ctl;
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
    step_ctl_reset(m.ctl);
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

// optional per-member accumulation of member-steps and of step trials (one
// per stepper call: a step plus its rejected trials), for benchmark accounting
CLODE_DEV void count_steps(const ClodeArgs &a, unsigned int i, unsigned int step, unsigned int trials)
{
    if (a.steps) {
        a.steps[i] += step;
        a.steps[(size_t)a.nPts + i] += trials;
    }
}

// the member's time span: the launch's, or its own when tspan_m is given
CLODE_DEV void member_tspan(const ClodeArgs &a, unsigned int i, timetype &t0, timetype &t1)
{
    if (a.tspan_m) {
        const double *tm = (const double *)a.tspan_m;
        t0 = (timetype)tm[i];
        t1 = (timetype)tm[(size_t)a.nPts + i];
    } else {
        t0 = (timetype)a.t0;
        t1 = (timetype)a.t1;
    }
}

#if CLODE_PREC == 3
#define CLODE_STEP(m)                                                                              \
    step_attempt(m.ti, m.xi, m.xc, m.dxi, m.p, sp, m.dt, t1, m.auxi, m.wi, m.rd, a.sp_reltol,           \
                 (realtype)(a.sp_abstol / a.sp_reltol), a.sp_dtmax, m.ctl)
#else
#define CLODE_STEP(m) step_attempt(m.ti, m.xi, m.dxi, m.p, sp, m.dt, t1, m.auxi, m.wi, m.rd, m.ctl)
#endif

// common start of every driver: load the member, draw the initial noise and
// evaluate the right-hand side at t0
CLODE_DEV void begin_member(Member &m, const ClodeArgs &a, unsigned int i, timetype t0)
{
    m.ti = t0;
    load_member(m, a, i);
    initial_noise(m);
    getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
}

struct TransientLane {
    Member m;
    timetype t1;
    unsigned int i, step, trials;
    CLODE_DEV void begin(const ClodeArgs &a, unsigned int idx)
    {
        i = idx;
        timetype t0;
        member_tspan(a, i, t0, t1);
        begin_member(m, a, i, t0);
        step = 0;
        trials = 0;
    }
    CLODE_DEV bool running(const SolverParamsR &sp) const { return m.ti <= t1 && step < sp.max_steps; }
    CLODE_DEV void advance(const ClodeArgs &a, const SolverParamsR &sp)
    {
        ++trials;
        if (CLODE_STEP(m))
            ++step;
    }
    CLODE_DEV void finish(const ClodeArgs &a)
    {
        store_member(m, a, i);
        count_steps(a, i, step, trials);
    }
};

#ifdef CLODE_OBSERVER
#ifndef CLODE_TWO_PASS
#define CLODE_TWO_PASS 0
#endif

// Warm-up record of a member's accepted steps (replay programs only; the
// layout is clode_replay_record's): record k holds the member after its
// (k+1)-th accepted step, everything the features pass's observer and, after
// the last record, its integration read from the member.
#define CLODE_REC_LEN (2 + 2 * N_VAR + N_AUX + (CLODE_PREC == 3 ? N_VAR : 0))
static_assert(CLODE_REC_LEN == clode_replay_record(N_VAR, N_AUX, CLODE_PREC), "warm-up record layout");

CLODE_DEV double *replay_record(const ClodeArgs &a, unsigned int i, unsigned int k)
{
    const size_t chunk = (size_t)(k / CLODE_REPLAY_CHUNK) * a.nPts + i;
    return a.replay + (chunk * CLODE_REPLAY_CHUNK + k % CLODE_REPLAY_CHUNK) * CLODE_REC_LEN;
}

CLODE_DEV void replay_write(const ClodeArgs &a, unsigned int i, unsigned int k, const Member &m)
{
    double *r = replay_record(a, i, k);
    r[0] = (double)m.ti;
#pragma unroll
    for (int j = 0; j < N_VAR; ++j) {
        r[1 + j] = (double)m.xi[j];
        r[1 + N_VAR + j] = (double)m.dxi[j];
    }
#pragma unroll
    for (int j = 0; j < N_AUX; ++j)
        r[1 + 2 * N_VAR + j] = (double)m.auxi[j];
    r[1 + 2 * N_VAR + N_AUX] = (double)m.dt;
#if CLODE_PREC == 3
#pragma unroll
    for (int j = 0; j < N_VAR; ++j)
        r[2 + 2 * N_VAR + N_AUX + j] = (double)m.xc[j];
#endif
}

// record k into registers (issued ahead of its use by the replay's ring)
CLODE_DEV void replay_load(const ClodeArgs &a, unsigned int i, unsigned int k, double r[CLODE_REC_LEN])
{
    const double *p = replay_record(a, i, k);
#pragma unroll
    for (int f = 0; f < CLODE_REC_LEN; ++f)
        r[f] = p[f];
}

CLODE_DEV void replay_take(const double r[CLODE_REC_LEN], Member &m)
{
    m.ti = (timetype)r[0];
#pragma unroll
    for (int j = 0; j < N_VAR; ++j) {
        m.xi[j] = (realtype)r[1 + j];
        m.dxi[j] = (realtype)r[1 + N_VAR + j];
    }
#pragma unroll
    for (int j = 0; j < N_AUX; ++j)
        m.auxi[j] = (realtype)r[1 + 2 * N_VAR + j];
    m.dt = (decltype(m.dt))r[1 + 2 * N_VAR + N_AUX];
#if CLODE_PREC == 3
#pragma unroll
    for (int j = 0; j < N_VAR; ++j)
        m.xc[j] = (realtype)r[2 + 2 * N_VAR + N_AUX + j];
#endif
}

// REC: also write the warm-up record (replay programs)
template <bool REC>
struct InitObsLaneT {
    Member m;
    Observer ob;
    ObserverParamsR op;
    timetype t0, t1;
    unsigned int i, step, trials;
    CLODE_DEV void begin(const ClodeArgs &a, unsigned int idx)
    {
        i = idx;
        op = make_observer_params(a);
        member_tspan(a, i, t0, t1);
        begin_member(m, a, i, t0);
        ObsIO io{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0};
        ob.init(m.ti, m.xi, m.dxi, m.auxi, op, io);
        step = 0;
        trials = 0;
    }
    CLODE_DEV bool running(const SolverParamsR &sp) const
    {
        return CLODE_TWO_PASS && m.ti < t1 && step < sp.max_steps;
    }
    CLODE_DEV void advance(const ClodeArgs &a, const SolverParamsR &sp)
    {
        ++trials;
        if (!CLODE_STEP(m))
            return;  // trial rejected: retried on the next iteration
        if (REC && step < a.replay_cap)
            replay_write(a, i, step, m);
        ++step;
        ob.warmup(m.ti, m.xi, m.dxi, m.auxi, op);
    }
    CLODE_DEV void finish(const ClodeArgs &a)
    {
        if (REC)
            a.replay_n[i] = step < a.replay_cap ? step : a.replay_cap;
        if (CLODE_TWO_PASS) {
            m.ti = t0;
            load_state(m, a, i);
            getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
            count_steps(a, i, step, trials);
        }
        ob.initEventDetector(m.ti, m.xi, m.dxi, m.auxi, op);
        ObsStore st{ObsIO{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0}};
        ob.fields(st);
    }
};
typedef InitObsLaneT<false> InitObsLane;

// REPLAY: the first replay_n[i] accepted steps are read back from the
// warm-up record instead of integrated.  The warm-up pass integrates the
// same member from the same state, step size and span with the same
// controller, and its loop (t < t1) ends no later than this one (t <= t1),
// so record k is exactly the member after this pass's (k+1)-th step; the
// observer sees the same values in the same order, and after the last record
// the member (state, derivative, aux, time and step size; the controller
// holds nothing else between steps) continues integrating where the warm-up
// record ends.  Only replayed steps are left out of the step counts.
// The records are loaded CLODE_REPLAY_DEPTH steps ahead of their use into a
// ring of registers (slot k mod depth holds record k), so that the loads of
// the next steps are in flight while the observer works on this one.
// Measured on the lactotroph (65,536 members, one wave per SIMD): depths 1,
// 2 and 4 replay in the same time (26.3-26.6 ms); the compiler's waits on
// the ring's loop-carried registers let no more than one step overlap, so
// the default is the one-step lookahead with the fewest registers.
#ifndef CLODE_REPLAY_DEPTH
#define CLODE_REPLAY_DEPTH 1
#endif
template <bool REPLAY>
struct FeaturesLaneT {
    Member m;
    Observer ob;
    ObserverParamsR op;
    timetype t0, t1;
    unsigned int i, step, trials, nrep;
    bool terminal;
    double ring[REPLAY ? CLODE_REPLAY_DEPTH : 1][CLODE_REC_LEN];
    CLODE_DEV void begin(const ClodeArgs &a, unsigned int idx)
    {
        i = idx;
        op = make_observer_params(a);
        member_tspan(a, i, t0, t1);
        m.ti = t0;
        load_member(m, a, i);
        {
            ObsLoad ld{ObsIO{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0}};
            ob.fields(ld);
        }
        initial_noise(m);
        getRHS((realtype)m.ti, m.xi, m.p, m.dxi, m.auxi, m.wi);
        step = 0;
        trials = 0;
        nrep = REPLAY ? a.replay_n[i] : 0u;
        terminal = false;
        if (REPLAY) {
#pragma unroll
            for (int d = 0; d < CLODE_REPLAY_DEPTH; ++d)
                if ((unsigned int)d < nrep)
                    replay_load(a, i, (unsigned int)d, ring[d]);
        }
    }
    CLODE_DEV bool running(const SolverParamsR &sp) const { return !terminal && m.ti <= t1 && step < sp.max_steps; }
    // one accepted step: the observer's update and event test
    CLODE_DEV void observe(const ClodeArgs &a)
    {
        const ObsIO io{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0};
        ++step;
        ob.update(m.ti, m.xi, m.dxi, m.auxi, op, io);
        if (ob.eventFunction(m.ti, m.xi, m.dxi, m.auxi, op))
            terminal = ob.computeEventFeatures(m.ti, m.xi, m.dxi, m.auxi, op, io);
    }
    CLODE_DEV void advance(const ClodeArgs &a, const SolverParamsR &sp)
    {
        if (REPLAY && step < nrep) {
            // the member's whole replay, in bursts of depth steps: slot d
            // holds record start + d (bursts start at multiples of the
            // depth), the unrolled slots keep the ring in registers, and
            // every slot's next load is issued (a valid record, the last one
            // near the end) so that each burst leaves the same loads in
            // flight
            bool go = true;
            while (go) {
#pragma unroll
                for (int d = 0; d < (REPLAY ? CLODE_REPLAY_DEPTH : 1); ++d) {
                    if (go) {
                        replay_take(ring[d], m);
                        const unsigned int ahead = step + CLODE_REPLAY_DEPTH;
                        replay_load(a, i, ahead < nrep ? ahead : nrep - 1u, ring[d]);
                        observe(a);
                        go = step < nrep && running(sp);
                    }
                }
            }
            return;
        }
        ++trials;
        if (!CLODE_STEP(m))
            return;  // trial rejected: retried on the next iteration
        observe(a);
    }
    CLODE_DEV void finish(const ClodeArgs &a)
    {
        const ObsIO io{(ostore *)a.odr, a.odu, a.nPts, i, 0, 0};
        FeatOut F{(ostore *)a.F, a.nPts, i, 0};
        ob.finalizeFeatures(m.ti, F, io);
        ob.finalizeObserverData(m.ti, t0);
        ObsStore st{io};
        ob.fields(st);
        store_member(m, a, i);
        count_steps(a, i, REPLAY ? step - (step < nrep ? step : nrep) : step, trials);
    }
};
typedef FeaturesLaneT<false> FeaturesLane;
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

struct TrajectoryLane {
    Member m;
    timetype t1;
    unsigned int i, step, storeix, trials;
    CLODE_DEV void begin(const ClodeArgs &a, unsigned int idx)
    {
        i = idx;
        timetype t0;
        member_tspan(a, i, t0, t1);
        begin_member(m, a, i, t0);
        storeix = 0;
        if (a.sp_max_store > 0)
            trajectory_store(a, i, storeix, m);
        step = 0;
        trials = 0;
    }
    CLODE_DEV bool running(const SolverParamsR &sp) const
    {
        return m.ti <= t1 && step < sp.max_steps && storeix < sp.max_store;
    }
    CLODE_DEV void advance(const ClodeArgs &a, const SolverParamsR &sp)
    {
        // Unlike the other drivers, finish the step here: the lanes of a wave
        // then store the same sample index together, so each store
        // instruction writes one contiguous run of the [slot][var][member]
        // buffers.  (Letting rejected lanes fall behind spreads a wave's
        // stores over many sample rows; measured 2.7x slower on 65,536
        // members x 1024 samples.)
        ++trials;
        while (!CLODE_STEP(m))
            ++trials;
        ++step;
        if (step % sp.nout == 0) {
            ++storeix;
            if (storeix < sp.max_store)
                trajectory_store(a, i, storeix, m);
        }
    }
    CLODE_DEV void finish(const ClodeArgs &a)
    {
        a.nStored[i] = (int)storeix;
        store_member(m, a, i);
        count_steps(a, i, step, trials);
    }
};

// one member per lane, to completion
template <class Lane>
CLODE_DEV void run_grid(const ClodeArgs &a)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nPts)
        return;
    const SolverParamsR sp = make_solver_params(a);
    Lane L;
    L.begin(a, i);
    while (L.running(sp))
        L.advance(a, sp);
    L.finish(a);
}

// Work queue: queue position q maps to member order[q * q_stride] (or
// q * q_stride without an order array); positions are claimed with one atomic
// add on work_counter, which the host zeroes before the launch.
CLODE_DEV bool claim_member(const ClodeArgs &a, unsigned int &idx)
{
    const unsigned int q = atomicAdd(a.work_counter, 1u);
    if (q >= a.q_count)
        return false;
    const unsigned int pos = q * a.q_stride;
    idx = a.q_order ? a.q_order[pos] : pos;
    return true;
}

// Persistent lanes: run up to q_slice iterations of the member's loop, then
// check; a lane whose member has finished stores it and claims the next one.
// Every iteration makes a step attempt for a member whose step count is
// bounded by max_steps (and attempts per step by CLODE_MAX_ATTEMPTS) or claims
// a queue position, so every lane leaves the loop once the queue is empty.
template <class Lane>
CLODE_DEV void run_queue(const ClodeArgs &a)
{
    const SolverParamsR sp = make_solver_params(a);
    const unsigned int slice = a.q_slice > 0 ? a.q_slice : 1u;
    Lane L;
    unsigned int idx = 0;
    bool have = claim_member(a, idx);
    if (have)
        L.begin(a, idx);
    while (have) {
        for (unsigned int k = 0; k < slice && L.running(sp); ++k)
            L.advance(a, sp);
        if (!L.running(sp)) {
            L.finish(a);
            have = claim_member(a, idx);
            if (have)
                L.begin(a, idx);
        }
    }
}

#ifndef CLODE_BATCH
extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_transient(const ClodeArgs a)
{
    run_grid<TransientLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_trajectory(const ClodeArgs a)
{
    run_grid<TrajectoryLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_queue_transient(const ClodeArgs a)
{
    run_queue<TransientLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_queue_trajectory(const ClodeArgs a)
{
    run_queue<TrajectoryLane>(a);
}

#ifdef CLODE_OBSERVER
extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_initialize_observer(const ClodeArgs a)
{
    run_grid<InitObsLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_features(const ClodeArgs a)
{
    run_grid<FeaturesLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_queue_initialize_observer(const ClodeArgs a)
{
    run_queue<InitObsLane>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_queue_features(const ClodeArgs a)
{
    run_queue<FeaturesLane>(a);
}

#if defined(CLODE_REPLAY) && CLODE_TWO_PASS
// Warm-up recording and features replay (grid launches): features(true) as
// these two launches gives the results of clode_initialize_observer and
// clode_features bit for bit, with the second pass's integration replaced by
// reading the first pass's record back from HBM.
extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_initialize_observer_record(const ClodeArgs a)
{
    run_grid<InitObsLaneT<true>>(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_features_replay(const ClodeArgs a)
{
    run_grid<FeaturesLaneT<true>>(a);
}
#endif

// transient -> x0 <- xf -> observer warm-up and initialisation -> features,
// member by member in one launch: the same lanes as the four separate calls,
// each phase reading what the member's previous phase stored, so the results
// are those of transient(), shift_x0(), features(true) with the two spans.
template <class Lane>
CLODE_DEV void run_member(const ClodeArgs &a, const SolverParamsR &sp, unsigned int i)
{
    Lane L;
    L.begin(a, i);
    while (L.running(sp))
        L.advance(a, sp);
    L.finish(a);
}

extern "C" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) CLODE_KERNEL_ATTRS clode_sequence(const ClodeSeqArgs s)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.tr.nPts)
        return;
    const SolverParamsR sp = make_solver_params(s.tr);
    run_member<TransientLane>(s.tr, sp, i);
    const size_t n = s.tr.nPts;
#pragma unroll
    for (int j = 0; j < N_VAR; ++j)  // shift_x0 for this member
        ((xstore *)s.tr.x0)[j * n + i] = ((const xstore *)s.tr.xf)[j * n + i];
    run_member<InitObsLane>(s.fe, sp, i);
    run_member<FeaturesLane>(s.fe, sp, i);
}

// The feature matrix F ([nF][nPts], member index fastest, as the observers
// write it) as double, member-major ([nPts][nF]): the layout the host API
// hands out, transposed here instead of on the host.  A block moves a tile
// of 64 members x 32 features through LDS, so that both the reads (along
// members) and the writes (along features) are contiguous runs.
extern "C" __global__ void __launch_bounds__(64) clode_f_members(const ostore *F, double *out, unsigned int nPts,
                                                                unsigned int nF)
{
    __shared__ double tile[32][65];
    const unsigned int m0 = blockIdx.x * 64u, f0 = blockIdx.y * 32u, t = threadIdx.x;
    for (unsigned int k = 0; k < 32u; ++k)
        if (f0 + k < nF && m0 + t < nPts)
            tile[k][t] = (double)F[(size_t)(f0 + k) * nPts + m0 + t];
    __syncthreads();
    for (unsigned int L = t; L < 64u * 32u; L += 64u) {
        const unsigned int r = L / 32u, c = L % 32u;
        if (m0 + r < nPts && f0 + c < nF)
            out[(size_t)(m0 + r) * nF + f0 + c] = tile[c][r];
    }
}
#endif  // CLODE_OBSERVER
#endif  // CLODE_BATCH

#endif  // CLODE_KERNELS_H
