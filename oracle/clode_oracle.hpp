// CPU oracle for the clODE ensemble-integration path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain restatement of the reference's
// algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg to verify GPU results (and as the timed CPU baseline).
// The library never includes, links or calls it.
//
// It follows the reference's own data layout and control flow rather than the
// GPU design: full ObserverData structs with 3-point buffers for every
// variable, run-time eVarIx/fVarIx and event-list length, one member at a time
// through the reference's loops.  Sources restated (paths relative to the
// clODE repository):
//   drivers   clode/cpp/transient.cl:9-77, initializeObserver.cl:9-84,
//             features.cl:10-106, trajectory.cl:14-112
//   steppers  clode/cpp/steppers/fixed_explicit_step.clh:9-36,
//             fixed_explicit_Euler.clh:5-17, fixed_explicit_Trapezoidal.clh:5-22,
//             fixed_explicit_RK4.clh:5-39, adaptive_explicit_step.clh:4-80,
//             adaptive_bs23.clh:3-62, adaptive_dp45.clh:4-110
//   RNG       clode/cpp/clODE_random.cl:93-183
//   observers clode/cpp/observers/observer_threshold_2.clh:121-470,
//             observer_local_maximum.clh:92-328, observer_neighborhood_2.clh:90-313
//   utilities clode/cpp/clODE_utilities.cl:31-177
//   host      clode/cpp/CLODE.cpp:174-242 (nPts change: dt refill, reseed),
//             :447-464 (seedRNG(int)), :495-515 (shifts),
//             clode/cpp/CLODEfeatures.cpp:143-259 (observer (re)initialisation)
//
// Elementary functions (exp/log/pow/pown in the RHS, the controller and the
// Gaussian generator) are the deterministic routines of
// clode_amd/csrc/device/cm_math.h, which the tests pin against glibc; sqrt,
// fma, nextafter, fmin/fmax are IEEE operations.  Compiled with clang++
// -ffp-contract=on -mfma, like the hipRTC programs, so multiply-adds are fused
// at the same places as in OpenCL C.
//
// Known reference undefined behaviour, resolved as the GPU path does:
// thresh2's local-extremum buffer has 3 entries (H7); localmax skips the
// local-minimum list store while eventcount == 0 (H8); the trajectory never
// stores sample index max_store (H9).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace oracle {

// Elementary operations are called unqualified, so that a real type other
// than float / double (the operation-counting type of flop_count.cpp) brings
// its own overloads by argument-dependent lookup.
using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::nextafter;
using std::sqrt;

// Hooks at the start of a step and at the end of every step trial (0
// rejected, 1 accepted, 2 failed at hmin), and a marker for values that vary
// during the integration; all empty except in the operation-counting build.
#ifndef CLODE_ORACLE_TRIAL_BEGIN
#define CLODE_ORACLE_TRIAL_BEGIN()
#endif
// step trials of the calling thread (one per stepper call, whatever its outcome)
inline unsigned long long &trial_count()
{
    static thread_local unsigned long long n = 0;
    return n;
}
#ifndef CLODE_ORACLE_TRIAL_END
#define CLODE_ORACLE_TRIAL_END(outcome) (++oracle::trial_count())
#endif
template <class R>
inline void mark_varying(R &)
{
}

enum { ST_EULER = 1, ST_HEUN = 2, ST_RK4 = 3, ST_BS23 = 4, ST_DOPRI5 = 5, ST_SEULER = 6, ST_HE12 = 7 };
enum { OB_NONE = 0, OB_THRESH2 = 1, OB_LOCALMAX = 2, OB_NHOOD2 = 3, OB_BASIC = 4, OB_BASICALL = 5, OB_NHOOD1 = 6 };

template <class R>
struct Consts;
template <>
struct Consts<float> {
    static float big() { return 3.40282347e+38f; }
    static float c(double, float f) { return f; }
};
template <>
struct Consts<double> {
    static double big() { return 1.7976931348623157e308; }
    static double c(double d, float) { return d; }
};
// RCONST(x): the literal in working precision (float literal folded in float)
#define OC(x) Consts<R>::c(x, x##f)

template <class R>
struct SolverParamsT {
    R dt, dtmax, abstol, reltol;
    unsigned int max_steps, max_store, nout;
};

template <class R>
struct ObserverParamsT {
    unsigned int eVarIx, fVarIx, maxEventCount, maxEventTimestamps;
    R minXamp, minIMI, nHoodRadius, xUpThresh, xDownThresh, dxUpThresh, dxDownThresh, eps_dx;
};

// ------------------------------------------------------------------- RNG
template <class R>
struct RngData {
    uint64_t state[2];
    bool randnUselast;
    R randnLast;
};

inline uint64_t rotl(const uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

inline uint64_t next(uint64_t s[2])
{
    const uint64_t s0 = s[0];
    uint64_t s1 = s[1];
    const uint64_t result = s0 + s1;
    s1 ^= s0;
    s[0] = rotl(s0, 55) ^ s1 ^ (s1 << 14);
    s[1] = rotl(s1, 36);
    return result;
}

template <class R>
inline R rand01(uint64_t s[2])
{
    uint64_t result = next(s);
    return (R)result * OC(5.421010862427522e-20);
}

template <class R>
inline R randn(RngData<R> &rd)
{
    R x1, x2, w, y1;
    if (rd.randnUselast) {
        y1 = rd.randnLast;
        rd.randnUselast = false;
    } else {
        int guard = 0;  // same bound as the library: only the all-zero state hits it
        do {
            x1 = OC(2.0) * rand01<R>(rd.state) - OC(1.0);
            x2 = OC(2.0) * rand01<R>(rd.state) - OC(1.0);
            w = x1 * x1 + x2 * x2;
        } while (w >= OC(1.0) && ++guard < 4096);
        w = sqrt((-OC(2.0) * cm::log(w)) / w);
        y1 = x1 * w;
        rd.randnLast = x2 * w;
        rd.randnUselast = true;
    }
    return y1;
}

// fixed-capacity scratch vector (no allocation in the step loop)
constexpr int MAXV = 64;
template <class R>
struct Buf {
    R v[MAXV];
    R *data() { return v; }
    R &operator[](int i) { return v[i]; }
};

// --------------------------------------------------------------- problem
// The right-hand side is supplied by the model translation unit as
// getRHS(t, x, p, dx, aux, w) overloads for float and double.
template <class R>
struct Problem {
    int nVar, nPar, nAux, nWiener;
    void (*rhs)(const R, const R *, const R *, R *, R *, const R *);
};

// ------------------------------------------------------------- steppers
template <class R>
struct Integrator {
    const Problem<R> &pb;
    int stepper;
    explicit Integrator(const Problem<R> &p, int s) : pb(p), stepper(s) {}

    void rhs(R t, const R *x, const R *p, R *dx, R *aux, const R *w) const { pb.rhs(t, x, p, dx, aux, w); }

    // fixed_explicit_Euler.clh:5-17
    void euler_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, const R *wi) const
    {
        for (int k = 0; k < pb.nVar; k++)
            xi[k] += dt * k1[k];
        *ti += dt;
    }
    // fixed_explicit_Trapezoidal.clh:5-22
    void heun_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, const R *wi) const
    {
        R th = *ti + dt;
        Buf<R> tmp, k2;
        for (int k = 0; k < pb.nVar; k++)
            tmp[k] = fma(dt, k1[k], xi[k]);
        rhs(th, tmp.data(), p, k2.data(), aux, wi);
        for (int k = 0; k < pb.nVar; k++)
            xi[k] += dt * OC(0.5) * (k1[k] + k2[k]);
        *ti = th;
    }
    // fixed_explicit_RK4.clh:5-39
    void rk4_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, const R *wi) const
    {
        const int N = pb.nVar;
        Buf<R> tmp, k2, k3, k4;
        R h2 = dt * OC(0.5);
        R th2 = *ti + h2;
        R th = *ti + dt;
        for (int k = 0; k < N; k++)
            tmp[k] = xi[k] + h2 * k1[k];
        rhs(th2, tmp.data(), p, k2.data(), aux, wi);
        for (int k = 0; k < N; k++)
            tmp[k] = xi[k] + h2 * k2[k];
        rhs(th2, tmp.data(), p, k3.data(), aux, wi);
        for (int k = 0; k < N; k++)
            tmp[k] = xi[k] + dt * k3[k];
        rhs(th, tmp.data(), p, k4.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xi[k] += dt * (k1[k] + OC(2.0) * k2[k] + OC(2.0) * k3[k] + k4[k]) / OC(6.0);
        *ti = th;
    }
    // fixed_explicit_step.clh:9-36
    int fixed_stepper(R *ti, R *xi, R *k1, const R *p, const SolverParamsT<R> &sp, R *dt, R t1, R *aux, R *wi,
                      RngData<R> &rd) const
    {
        CLODE_ORACLE_TRIAL_BEGIN();
        if (stepper == ST_EULER || stepper == ST_SEULER)
            euler_step(ti, xi, k1, p, *dt, aux, wi);
        else if (stepper == ST_HEUN)
            heun_step(ti, xi, k1, p, *dt, aux, wi);
        else
            rk4_step(ti, xi, k1, p, *dt, aux, wi);
        if (stepper == ST_SEULER)
            for (int j = 0; j < pb.nWiener; ++j)
                wi[j] = randn(rd) / sqrt(*dt);
        rhs(*ti, xi, p, k1, aux, wi);
        CLODE_ORACLE_TRIAL_END(1);
        return 0;
    }

    // adaptive_he12.clh:9-40 (Heun-Euler 1(2): an Euler predictor, the
    // trapezoidal average as the new state, half the slope difference as the
    // error; the slope at the new state is evaluated again for k1)
    R he12_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, R *err, const R *wi) const
    {
        const int N = pb.nVar;
        R tNew = *ti + dt;
        R newDt = tNew - *ti;
        Buf<R> xtmp, k2;
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * k1[k];
        rhs(tNew, xtmp.data(), p, k2.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xi[k] = xi[k] + newDt * OC(0.5) * (k1[k] + k2[k]);
        for (int k = 0; k < N; k++)
            err[k] = newDt * OC(0.5) * (-k1[k] + k2[k]);
        rhs(tNew, xi, p, k2.data(), aux, wi);
        for (int k = 0; k < N; k++)
            k1[k] = k2[k];
        *ti = tNew;
        return newDt;
    }

    // adaptive_bs23.clh:3-62
    R bs23_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, R *err, const R *wi) const
    {
        const int N = pb.nVar;
        const R B1 = OC(2.0) / OC(9.0), B2 = OC(1.0) / OC(3.0), B3 = OC(4.0) / OC(9.0);
        const R E1 = OC(-5.0) / OC(72.0), E2 = OC(1.0) / OC(12.0), E3 = OC(1.0) / OC(9.0),
                E4 = OC(-1.0) / OC(8.0);
        R tNew = *ti + dt;
        R newDt = tNew - *ti;
        Buf<R> xtmp, k2, k3, k4;
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * OC(0.5) * k1[k];
        rhs(*ti + newDt * OC(0.5), xtmp.data(), p, k2.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * OC(0.75) * k2[k];
        rhs(*ti + newDt * OC(0.75), xtmp.data(), p, k3.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xi[k] = xi[k] + newDt * (B1 * k1[k] + B2 * k2[k] + B3 * k3[k]);
        rhs(tNew, xi, p, k4.data(), aux, wi);
        for (int k = 0; k < N; k++) {
            err[k] = newDt * (E1 * k1[k] + E2 * k2[k] + E3 * k3[k] + E4 * k4[k]);
            k1[k] = k4[k];
        }
        *ti = tNew;
        return newDt;
    }

    // adaptive_dp45.clh:4-110
    R dopri5_step(R *ti, R *xi, R *k1, const R *p, R dt, R *aux, R *err, const R *wi) const
    {
        const int N = pb.nVar;
        const R A2 = OC(1.0) / OC(5.0), A3 = OC(3.0) / OC(10.0), A4 = OC(4.0) / OC(5.0), A5 = OC(8.0) / OC(9.0);
        const R B21 = OC(1.0) / OC(5.0);
        const R B31 = OC(3.0) / OC(40.0), B32 = OC(9.0) / OC(40.0);
        const R B41 = OC(44.0) / OC(45.0), B42 = OC(-56.0) / OC(15.0), B43 = OC(32.0) / OC(9.0);
        const R B51 = OC(19372.0) / OC(6561.0), B52 = OC(-25360.0) / OC(2187.0), B53 = OC(64448.0) / OC(6561.0),
                B54 = OC(-212.0) / OC(729.0);
        const R B61 = OC(9017.0) / OC(3168.0), B62 = OC(-355.0) / OC(33.0), B63 = OC(46732.0) / OC(5247.0),
                B64 = OC(49.0) / OC(176.0), B65 = OC(-5103.0) / OC(18656.0);
        const R C1 = OC(35.0) / OC(384.0), C3 = OC(500.0) / OC(1113.0), C4 = OC(125.0) / OC(192.0),
                C5 = OC(-2187.0) / OC(6784.0), C6 = OC(11.0) / OC(84.0);
        const R E1 = OC(71.0) / OC(57600.0), E3 = OC(-71.0) / OC(16695.0), E4 = OC(71.0) / OC(1920.0),
                E5 = OC(-17253.0) / OC(339200.0), E6 = OC(22.0) / OC(525.0), E7 = OC(-1.0) / OC(40.0);
        R tNew = *ti + dt;
        R newDt = tNew - *ti;
        Buf<R> xtmp, k2, k3, k4, k5, k6, k7;
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * (B21 * k1[k]);
        rhs(*ti + A2 * newDt, xtmp.data(), p, k2.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * (B31 * k1[k] + B32 * k2[k]);
        rhs(*ti + A3 * newDt, xtmp.data(), p, k3.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * (B41 * k1[k] + B42 * k2[k] + B43 * k3[k]);
        rhs(*ti + A4 * newDt, xtmp.data(), p, k4.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * (B51 * k1[k] + B52 * k2[k] + B53 * k3[k] + B54 * k4[k]);
        rhs(*ti + A5 * newDt, xtmp.data(), p, k5.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xtmp[k] = xi[k] + newDt * (B61 * k1[k] + B62 * k2[k] + B63 * k3[k] + B64 * k4[k] + B65 * k5[k]);
        rhs(*ti + newDt, xtmp.data(), p, k6.data(), aux, wi);
        for (int k = 0; k < N; k++)
            xi[k] = xi[k] + newDt * (C1 * k1[k] + C3 * k3[k] + C4 * k4[k] + C5 * k5[k] + C6 * k6[k]);
        rhs(tNew, xi, p, k7.data(), aux, wi);
        for (int k = 0; k < N; k++) {
            err[k] = newDt * (E1 * k1[k] + E3 * k3[k] + E4 * k4[k] + E5 * k5[k] + E6 * k6[k] + E7 * k7[k]);
            k1[k] = k7[k];
        }
        *ti = tNew;
        return newDt;
    }

    // adaptive_explicit_step.clh:4-80
    int adaptive_stepper(R *ti, R *xi, R *k1, const R *p, const SolverParamsT<R> &sp, R *dt, R t1, R *aux, R *wi,
                         RngData<R> &rd) const
    {
        CLODE_ORACLE_TRIAL_BEGIN();
        const int N = pb.nVar;
        const bool bs = stepper == ST_BS23, he = stepper == ST_HE12;
        const R maxShrink = (bs || he) ? OC(0.5) : OC(0.1);
        const R maxGrow = OC(5.0);
        // the reference's pow(rtol/err, 1/(p+1)) as the fixed root x^(1/(p+1)),
        // p = 1 (he12: the square root), 2 (bs23) or 4 (dopri5), shared with
        // the device (cm_math.h)
        auto root = [bs, he](R v) { return he ? cm::root<2>(v) : bs ? cm::root<3>(v) : cm::root<5>(v); };
        R tNew, normErr;
        Buf<R> err, newxi, newk1;
        R newDt = *dt;
        R threshold = sp.abstol / sp.reltol;
        R hmin = OC(16.0) * fabs(fabs(nextafter(*ti, OC(1.1) * t1)) - *ti);
        bool noFailedSteps = true;
        // The reference's loop ends at hmin.  A rejected trial is followed by
        // a strictly smaller clamped step until the clamp reaches hmin, so a
        // trial equal to the previous one happens only at hmin when t + hmin
        // rounds to a longer step; the reference would repeat it forever, and
        // it fails at hmin here (stall).  4096 trials bound the loop in any
        // case; a sequence that ends in the reference needs at most ~2300.
        R lastTrial = OC(0.0);
        int attempts = 0;
        while (true) {
            newDt = fmin(fmax(newDt, hmin), sp.dtmax);  // OpenCL clamp
            if ((!noFailedSteps && newDt == lastTrial) || ++attempts > 4096) {
                *dt = hmin;
                CLODE_ORACLE_TRIAL_END(2);
                return -1;
            }
            lastTrial = newDt;
            tNew = *ti;
            for (int j = 0; j < N; j++) {
                newxi[j] = xi[j];
                newk1[j] = k1[j];
            }
            if (he)
                newDt = he12_step(&tNew, newxi.data(), newk1.data(), p, newDt, aux, err.data(), wi);
            else if (bs)
                newDt = bs23_step(&tNew, newxi.data(), newk1.data(), p, newDt, aux, err.data(), wi);
            else
                newDt = dopri5_step(&tNew, newxi.data(), newk1.data(), p, newDt, aux, err.data(), wi);
            for (int j = 0; j < N; j++)
                err[j] /= fmax(fmax(fabs(xi[j]), fabs(newxi[j])), threshold);
            normErr = OC(0.0);
            for (int k = 0; k < N; k++)
                normErr = fmax(fabs(err[k]), normErr);
            if (normErr > sp.reltol) {
                if (newDt <= hmin) {
                    *dt = hmin;
                    CLODE_ORACLE_TRIAL_END(2);
                    return -1;
                }
                if (noFailedSteps) {
                    noFailedSteps = false;
                    newDt *= fmax(maxShrink, OC(0.8) * root(sp.reltol / normErr));
                } else {
                    newDt *= OC(0.5);
                }
                CLODE_ORACLE_TRIAL_END(0);
            } else {
                break;
            }
        }
        if (noFailedSteps)
            newDt *= fmin(maxGrow, OC(0.8) * root(sp.reltol / normErr));
        newDt = fmin(newDt, t1 - tNew);
        newDt = fmin(fmax(newDt, hmin), sp.dtmax);
        *dt = newDt;
        *ti = tNew;
        for (int j = 0; j < N; j++) {
            xi[j] = newxi[j];
            k1[j] = newk1[j];
        }
        CLODE_ORACLE_TRIAL_END(1);
        return 0;
    }

    int step(R *ti, R *xi, R *k1, const R *p, const SolverParamsT<R> &sp, R *dt, R t1, R *aux, R *wi,
             RngData<R> &rd) const
    {
        if (stepper == ST_BS23 || stepper == ST_DOPRI5 || stepper == ST_HE12)
            return adaptive_stepper(ti, xi, k1, p, sp, dt, t1, aux, wi, rd);
        return fixed_stepper(ti, xi, k1, p, sp, dt, t1, aux, wi, rd);
    }
};

// -------------------------------------------------------------- utilities
template <class R>
inline void runningMean(R *mean, R newValue, unsigned int eventCount)
{
    if (eventCount == 1)
        *mean = newValue;
    else if (eventCount > 1)
        *mean += (newValue - *mean) / (R)eventCount;
}

template <class R>
inline R runningMeanTime(R mean, R newValue, R dt, R total_delta)
{
    return mean + (newValue - mean) * dt / total_delta;
}

// ---------------------------------------------------------- observer data
// Union of the fields of the observers' ObserverData structs
// (observer_threshold_2.clh:63-116, observer_local_maximum.clh:54-87,
// observer_neighborhood_2.clh:50-85, observer_basic.clh:19-29,
// observer_basic_allVar.clh:33-47, observer_neighborhood_1.clh:43-84); each
// observer touches its own subset.
template <class R>
struct ObserverData {
    R tbuffer[3];
    std::vector<R> xbuffer, dxbuffer;  // [3*N]
    std::vector<R> xTrajectoryMax, xTrajectoryMin, xTrajectoryMean, dxTrajectoryMax, dxTrajectoryMin;
    std::vector<R> xTrajectoryRange, x0;
    std::vector<R> auxTrajectoryMax, auxTrajectoryMin, auxTrajectoryMean;
    std::vector<R> list[4];  // thresh2: tUp, tDown; localmax: tMax, xMax, tMin, xMin; nhood2: tExit
    R nMaxima[3], period[3], upDuration[3], downDuration[3], duty[3], activeDip[3], stepDt[3];
    R IMI[3], amp[3];
    R fVarDownstateMean;
    R xGlobalMax, xGlobalMin, dxGlobalMax, dxGlobalMin, xUp, xDown, dxUp, dxDown;
    R t_start, t_last, tLastEvent, tThisDown, tLastMax, tLastMin, xLastMin, xThreshold;
    unsigned int thisNMaxima, stepcount, eventcount, inUpstate, foundX0, isInNhood;

    void resize(int N, int A, int S)
    {
        xbuffer.assign(3 * N, R(0));
        dxbuffer.assign(3 * N, R(0));
        for (auto *v : {&xTrajectoryMax, &xTrajectoryMin, &xTrajectoryMean, &dxTrajectoryMax, &dxTrajectoryMin,
                        &xTrajectoryRange, &x0})
            v->assign(N, R(0));
        for (auto *v : {&auxTrajectoryMax, &auxTrajectoryMin, &auxTrajectoryMean})
            v->assign(A, R(0));
        for (auto &l : list)
            l.assign(S, R(0));
    }
};

template <class R>
struct Observer {
    int kind;
    int N, A, S;
    ObserverParamsT<R> op;
    R big() const { return Consts<R>::big(); }

    // ---------------- init (thresh2 :121-194, localmax :92-134, nhood2 :90-138)
    void init(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        od.tbuffer[2] = ti;
        for (int j = 0; j < N; ++j) {
            od.xbuffer[j * 3 + 2] = xi[j];
            od.dxbuffer[j * 3 + 2] = dxi[j];
        }
        for (int j = 0; j < N; ++j) {
            od.xTrajectoryMean[j] = 0;
            od.xTrajectoryMax[j] = -big();
            od.xTrajectoryMin[j] = big();
            od.dxTrajectoryMax[j] = -big();
            od.dxTrajectoryMin[j] = big();
            od.xTrajectoryRange[j] = 0;
            od.x0[j] = xi[j];
        }
        for (int j = 0; j < A; ++j) {
            od.auxTrajectoryMax[j] = -big();
            od.auxTrajectoryMin[j] = big();
            od.auxTrajectoryMean[j] = 0;
        }
        for (auto &l : od.list)
            std::fill(l.begin(), l.end(), R(0));
        for (R *s : {od.nMaxima, od.period, od.upDuration, od.downDuration, od.duty, od.activeDip, od.stepDt,
                     od.IMI, od.amp}) {
            s[0] = -big();
            s[1] = big();
            s[2] = 0;
        }
        od.fVarDownstateMean = 0;
        od.xGlobalMax = -big();
        od.xGlobalMin = big();
        od.dxGlobalMax = -big();
        od.dxGlobalMin = big();
        od.xUp = od.xDown = od.dxUp = od.dxDown = 0;
        od.t_start = ti;
        od.t_last = ti;
        od.tLastEvent = 0;
        od.tThisDown = 0;
        od.tLastMax = 0;
        od.tLastMin = 0;
        od.xLastMin = big();
        od.xThreshold = 0;
        od.inUpstate = 0;
        od.thisNMaxima = 0;
        od.stepcount = 0;
        od.eventcount = 0;
        od.foundX0 = 0;
        od.isInNhood = 0;
    }

    void warmup(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        if (kind == OB_THRESH2) {  // :197-203
            od.xGlobalMax = fmax(od.xGlobalMax, xi[op.eVarIx]);
            od.xGlobalMin = fmin(od.xGlobalMin, xi[op.eVarIx]);
            od.dxGlobalMax = fmax(od.dxGlobalMax, dxi[op.eVarIx]);
            od.dxGlobalMin = fmin(od.dxGlobalMin, dxi[op.eVarIx]);
        } else if (kind == OB_NHOOD2) {  // :142-147
            for (int j = 0; j < N; ++j) {
                od.xTrajectoryMax[j] = fmax(xi[j], od.xTrajectoryMax[j]);
                od.xTrajectoryMin[j] = fmin(xi[j], od.xTrajectoryMin[j]);
            }
        }
    }

    void initEventDetector(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        if (kind == OB_THRESH2) {  // :206-224
            R xTrajectoryAmp = od.xGlobalMax - od.xGlobalMin;
            od.xUp = od.xGlobalMin + op.xUpThresh * xTrajectoryAmp;
            if (op.xDownThresh > 0)
                od.xDown = od.xGlobalMin + op.xDownThresh * xTrajectoryAmp;
            else
                od.xDown = od.xUp;
            od.dxUp = op.dxUpThresh * od.dxGlobalMax;
            if (op.dxDownThresh > 0)
                od.dxDown = op.dxDownThresh * od.dxGlobalMin;
            else
                od.dxDown = od.dxGlobalMin;
            od.inUpstate = xi[op.eVarIx] > od.xUp ? 1 : 0;
        } else if (kind == OB_NHOOD2) {  // :150-154
            for (int j = 0; j < N; ++j)
                od.xTrajectoryRange[j] = od.xTrajectoryMax[j] - od.xTrajectoryMin[j];
            od.xThreshold = od.xTrajectoryMin[op.eVarIx] + op.xDownThresh * od.xTrajectoryRange[op.eVarIx];
        }
    }

    // shared per-step extents and time means
    void extents(R ti, const R *xi, const R *dxi, const R *auxi, ObserverData<R> &od, R dt, R elapsedTime) const
    {
        for (int j = 0; j < N; ++j) {
            od.xTrajectoryMax[j] = fmax(xi[j], od.xTrajectoryMax[j]);
            od.xTrajectoryMin[j] = fmin(xi[j], od.xTrajectoryMin[j]);
            od.xTrajectoryMean[j] = runningMeanTime(od.xTrajectoryMean[j], xi[j], dt, elapsedTime);
            od.dxTrajectoryMax[j] = fmax(dxi[j], od.dxTrajectoryMax[j]);
            od.dxTrajectoryMin[j] = fmin(dxi[j], od.dxTrajectoryMin[j]);
        }
        for (int j = 0; j < A; ++j) {
            od.auxTrajectoryMax[j] = fmax(auxi[j], od.auxTrajectoryMax[j]);
            od.auxTrajectoryMin[j] = fmin(auxi[j], od.auxTrajectoryMin[j]);
            od.auxTrajectoryMean[j] = runningMeanTime(od.auxTrajectoryMean[j], auxi[j], dt, elapsedTime);
        }
    }

    void advanceBuffers(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        od.tbuffer[0] = od.tbuffer[1];
        od.tbuffer[1] = od.tbuffer[2];
        od.tbuffer[2] = ti;
        for (int j = 0; j < N; ++j) {
            od.xbuffer[j * 3 + 0] = od.xbuffer[j * 3 + 1];
            od.xbuffer[j * 3 + 1] = od.xbuffer[j * 3 + 2];
            od.xbuffer[j * 3 + 2] = xi[j];
            od.dxbuffer[j * 3 + 0] = od.dxbuffer[j * 3 + 1];
            od.dxbuffer[j * 3 + 1] = od.dxbuffer[j * 3 + 2];
            od.dxbuffer[j * 3 + 2] = dxi[j];
        }
    }

    void update(R ti, const R *xi, const R *dxi, const R *auxi, ObserverData<R> &od) const
    {
        ++od.stepcount;
        advanceBuffers(ti, xi, dxi, od);
        R dt = od.tbuffer[2] - od.tbuffer[1];
        R elapsedTime = ti - od.t_start;
        const unsigned f = op.fVarIx, e = op.eVarIx;
        if (kind == OB_THRESH2) {  // :308-407
            od.stepDt[0] = fmax(dt, od.stepDt[0]);
            od.stepDt[1] = fmin(dt, od.stepDt[1]);
            runningMean(&od.stepDt[2], dt, od.stepcount);
            extents(ti, xi, dxi, auxi, od, dt, elapsedTime);
            if (od.stepcount > 1) {
                if (od.dxbuffer[f * 3 + 1] > 0.0 && od.dxbuffer[f * 3 + 2] < 0.0) {
                    R buf[3];
                    for (int j = 0; j < 3; ++j)
                        buf[j] = od.xbuffer[f * 3 + j];
                    R mx = -big();
                    int ix = 0;
                    for (int k = 0; k < 3; k++)
                        if (buf[k] > mx) {
                            mx = buf[k];
                            ix = k;
                        }
                    R tThisMax = od.tbuffer[ix];
                    od.thisNMaxima++;
                    od.tLastMax = tThisMax;
                }
                if (od.dxbuffer[f * 3 + 1] < 0.0 && od.dxbuffer[f * 3 + 2] > 0.0) {
                    R buf[3];
                    for (int j = 0; j < 3; ++j)
                        buf[j] = od.xbuffer[f * 3 + j];
                    R mn = big();
                    int ix = 0;
                    for (int k = 0; k < 3; k++)
                        if (buf[k] < mn) {
                            mn = buf[k];
                            ix = k;
                        }
                    od.xLastMin = mn;
                    od.tLastMin = od.tbuffer[ix];
                }
                if (od.inUpstate) {
                    if (xi[e] <= od.xDown && dxi[e] >= od.dxDown) {
                        od.tThisDown = ti;
                        od.inUpstate = 0;
                        if (od.eventcount > 0 && od.eventcount <= (unsigned)S)
                            od.list[1][od.eventcount - 1] = od.tThisDown;
                        od.fVarDownstateMean = xi[f];
                    }
                } else {
                    R elapsedDownTime = (ti - od.tThisDown);
                    if (elapsedDownTime > 0.0)
                        od.fVarDownstateMean = runningMeanTime(od.fVarDownstateMean, xi[f], dt, elapsedDownTime);
                }
            }
        } else if (kind == OB_LOCALMAX) {  // :211-283
            extents(ti, xi, dxi, auxi, od, dt, elapsedTime);
            if (od.stepcount < 2)
                return;
            if (od.dxbuffer[f * 3 + 1] < 0.0 && od.dxbuffer[f * 3 + 2] > 0.0) {
                R buf[3];
                for (int j = 0; j < 3; ++j)
                    buf[j] = od.xbuffer[f * 3 + j];
                R mn = big();
                int ix = 0;
                for (int k = 0; k < 3; k++)
                    if (buf[k] < mn) {
                        mn = buf[k];
                        ix = k;
                    }
                od.tLastMin = od.tbuffer[ix];
                od.xLastMin = buf[ix];
                if (od.eventcount > 0 && od.eventcount <= (unsigned)S) {
                    od.list[2][od.eventcount - 1] = od.tLastMin;
                    od.list[3][od.eventcount - 1] = od.xLastMin;
                }
            }
        } else if (kind == OB_BASIC || kind == OB_BASICALL) {
            // observer_basic.clh:63-73, observer_basic_allVar.clh:91-111: time
            // means over the step since the previous update (t_last)
            R dtb = ti - od.t_last;
            od.t_last = ti;
            R el = ti - od.t_start;
            for (int j = 0; j < N; ++j) {
                if (kind == OB_BASIC && j != (int)f)
                    continue;
                od.xTrajectoryMax[j] = fmax(xi[j], od.xTrajectoryMax[j]);
                od.xTrajectoryMin[j] = fmin(xi[j], od.xTrajectoryMin[j]);
                od.xTrajectoryMean[j] = runningMeanTime(od.xTrajectoryMean[j], xi[j], dtb, el);
                od.dxTrajectoryMax[j] = fmax(dxi[j], od.dxTrajectoryMax[j]);
                od.dxTrajectoryMin[j] = fmin(dxi[j], od.dxTrajectoryMin[j]);
            }
            if (kind == OB_BASICALL)
                for (int j = 0; j < A; ++j) {
                    od.auxTrajectoryMax[j] = fmax(auxi[j], od.auxTrajectoryMax[j]);
                    od.auxTrajectoryMin[j] = fmin(auxi[j], od.auxTrajectoryMin[j]);
                    od.auxTrajectoryMean[j] = runningMeanTime(od.auxTrajectoryMean[j], auxi[j], dtb, el);
                }
        } else if (kind == OB_NHOOD1) {  // observer_neighborhood_1.clh:203-276
            od.stepDt[0] = fmax(dt, od.stepDt[0]);
            od.stepDt[1] = fmin(dt, od.stepDt[1]);
            runningMean(&od.stepDt[2], dt, od.stepcount);
            // step-count means, not time means
            for (int j = 0; j < N; ++j) {
                od.xTrajectoryMax[j] = fmax(xi[j], od.xTrajectoryMax[j]);
                od.xTrajectoryMin[j] = fmin(xi[j], od.xTrajectoryMin[j]);
                runningMean(&od.xTrajectoryMean[j], xi[j], od.stepcount);
                od.dxTrajectoryMax[j] = fmax(dxi[j], od.dxTrajectoryMax[j]);
                od.dxTrajectoryMin[j] = fmin(dxi[j], od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                od.auxTrajectoryMax[j] = fmax(auxi[j], od.auxTrajectoryMax[j]);
                od.auxTrajectoryMin[j] = fmin(auxi[j], od.auxTrajectoryMin[j]);
                runningMean(&od.auxTrajectoryMean[j], auxi[j], od.stepcount);
            }
            if (od.stepcount > 1) {
                if (!od.foundX0) {
                    // first local minimum of the event variable is the centre
                    if (od.dxbuffer[e * 3 + 1] <= 0.0 && od.dxbuffer[e * 3 + 2] > 0.0) {
                        R mn = big();
                        int ix = 0;
                        for (int k = 0; k < 3; k++)
                            if (od.xbuffer[e * 3 + k] < mn) {
                                mn = od.xbuffer[e * 3 + k];
                                ix = k;
                            }
                        od.tLastEvent = od.tbuffer[ix];
                        od.foundX0 = 1;
                        for (int j = 0; j < N; ++j)
                            od.x0[j] = od.xbuffer[j * 3 + ix];
                    }
                } else if (od.dxbuffer[f * 3 + 1] >= 0.0 && od.dxbuffer[f * 3 + 2] < 0.0) {
                    od.thisNMaxima++;
                }
            }
        } else if (kind == OB_NHOOD2) {  // :208-268
            od.stepDt[0] = fmax(dt, od.stepDt[0]);
            od.stepDt[1] = fmin(dt, od.stepDt[1]);
            runningMean(&od.stepDt[2], dt, od.stepcount);
            extents(ti, xi, dxi, auxi, od, dt, elapsedTime);
            if (od.stepcount < 2)
                return;
            if (od.foundX0) {
                if (od.dxbuffer[f * 3 + 1] >= 0.0 && od.dxbuffer[f * 3 + 2] < 0.0)
                    od.thisNMaxima++;
                return;
            }
            if (od.xbuffer[e * 3 + 1] > od.xThreshold && od.xbuffer[e * 3 + 2] < od.xThreshold) {
                od.foundX0 = 1;
                od.isInNhood = 1;
                for (int j = 0; j < N; ++j)
                    od.x0[j] = xi[j];
            }
        }
    }

    bool eventFunction(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        const unsigned f = op.fVarIx, e = op.eVarIx;
        if (kind == OB_THRESH2) {  // :227-240
            if (od.stepcount < 2)
                return false;
            if (od.xGlobalMax - od.xGlobalMin < op.minXamp)
                return false;
            if (od.inUpstate)
                return false;
            return (xi[e] > od.xUp && dxi[e] > od.dxUp);
        } else if (kind == OB_LOCALMAX) {  // :145-152
            if (od.stepcount < 2)
                return false;
            return (od.dxbuffer[f * 3 + 1] > 0.0 && od.dxbuffer[f * 3 + 2] < 0.0);
        } else if (kind == OB_BASIC || kind == OB_BASICALL) {
            return false;
        } else if (kind == OB_NHOOD1) {  // observer_neighborhood_1.clh:148-168
            if (od.stepcount < 2 || !od.foundX0)
                return false;
            if (od.xTrajectoryMax[f] - od.xTrajectoryMin[f] < op.minXamp)
                return false;
            unsigned int lastInNhood = od.isInNhood;
            Buf<R> d;
            for (int j = 0; j < N; ++j)
                d[j] = fabs(xi[j] - od.x0[j]) / (od.xTrajectoryMax[j] - od.xTrajectoryMin[j]);
            R result = 0;
            for (int k = 0; k < N; k++)
                result += d[k] * d[k];
            od.isInNhood = sqrt(result) <= op.nHoodRadius;
            return od.isInNhood && !lastInNhood;
        } else {  // nhood2 :157-175
            if (od.stepcount < 2 || !od.foundX0)
                return false;
            Buf<R> d;
            for (int j = 0; j < N; ++j)
                d[j] = (xi[j] - od.x0[j]) / od.xTrajectoryRange[j];
            unsigned int lastInNhood = od.isInNhood;
            R result = 0;
            for (int k = 0; k < N; k++)
                result += d[k] * d[k];
            od.isInNhood = sqrt(result) < op.nHoodRadius;
            return lastInNhood && !od.isInNhood;
        }
    }

    bool computeEventFeatures(R ti, const R *xi, const R *dxi, ObserverData<R> &od) const
    {
        if (kind == OB_THRESH2) {  // :243-305
            R tThisEvent;
            ++od.eventcount;
            od.inUpstate = 1;
            tThisEvent = ti;
            if (od.eventcount > 1) {
                int nPeriods = od.eventcount - 1;
                od.nMaxima[0] = fmax((R)od.thisNMaxima, od.nMaxima[0]);
                od.nMaxima[1] = fmin((R)od.thisNMaxima, od.nMaxima[1]);
                runningMean(&od.nMaxima[2], (R)od.thisNMaxima, nPeriods);
                R thisPeriod = tThisEvent - od.tLastEvent;
                od.period[0] = fmax(thisPeriod, od.period[0]);
                od.period[1] = fmin(thisPeriod, od.period[1]);
                runningMean(&od.period[2], thisPeriod, nPeriods);
                R thisUpDuration = od.tThisDown - od.tLastEvent;
                od.upDuration[0] = fmax(thisUpDuration, od.upDuration[0]);
                od.upDuration[1] = fmin(thisUpDuration, od.upDuration[1]);
                runningMean(&od.upDuration[2], thisUpDuration, nPeriods);
                R thisDownDuration = tThisEvent - od.tThisDown;
                od.downDuration[0] = fmax(thisDownDuration, od.downDuration[0]);
                od.downDuration[1] = fmin(thisDownDuration, od.downDuration[1]);
                runningMean(&od.downDuration[2], thisDownDuration, nPeriods);
                R thisduty = thisUpDuration / thisPeriod;
                od.duty[0] = fmax(thisduty, od.duty[0]);
                od.duty[1] = fmin(thisduty, od.duty[1]);
                runningMean(&od.duty[2], thisduty, nPeriods);
                R thisActiveDip = od.fVarDownstateMean - od.xLastMin;
                od.activeDip[0] = fmax(thisActiveDip, od.activeDip[0]);
                od.activeDip[1] = fmin(thisActiveDip, od.activeDip[1]);
                runningMean(&od.activeDip[2], thisActiveDip, nPeriods);
            }
            if (od.eventcount <= (unsigned)S)
                od.list[0][od.eventcount - 1] = tThisEvent;
            od.tLastEvent = tThisEvent;
            od.thisNMaxima = 0;
            return od.eventcount == op.maxEventCount;
        } else if (kind == OB_LOCALMAX) {  // :156-205
            const unsigned f = op.fVarIx;
            R buf[3];
            for (int j = 0; j < 3; ++j)
                buf[j] = od.xbuffer[f * 3 + j];
            R mx = -big();
            int ix = 0;
            for (int k = 0; k < 3; k++)
                if (buf[k] > mx) {
                    mx = buf[k];
                    ix = k;
                }
            R tThisMax = od.tbuffer[ix];
            R xThisMax = buf[ix];
            ++od.eventcount;
            if (od.eventcount > 1) {
                R thisIMI = tThisMax - od.tLastMax;
                od.IMI[0] = fmax(thisIMI, od.IMI[0]);
                od.IMI[1] = fmin(thisIMI, od.IMI[1]);
                runningMean(&od.IMI[2], thisIMI, od.eventcount - 1);
                R thisAmp = xThisMax - od.xLastMin;
                od.amp[0] = fmax(thisAmp, od.amp[0]);
                od.amp[1] = fmin(thisAmp, od.amp[1]);
                runningMean(&od.amp[2], thisAmp, od.eventcount - 1);
            }
            od.tLastMax = tThisMax;
            if (od.eventcount <= (unsigned)S) {
                od.list[0][od.eventcount - 1] = tThisMax;
                od.list[1][od.eventcount - 1] = xThisMax;
            }
            return od.eventcount == op.maxEventCount;
        } else if (kind == OB_BASIC || kind == OB_BASICALL) {
            return false;
        } else if (kind == OB_NHOOD1) {  // observer_neighborhood_1.clh:172-198
            ++od.eventcount;
            R tThisEvent = ti;
            if (od.eventcount > 1) {
                od.nMaxima[0] = fmax((R)od.thisNMaxima, od.nMaxima[0]);
                od.nMaxima[1] = fmin((R)od.thisNMaxima, od.nMaxima[1]);
                runningMean(&od.nMaxima[2], (R)od.thisNMaxima, od.eventcount - 1);
                R thisPeriod = tThisEvent - od.tLastEvent;
                od.period[0] = fmax(thisPeriod, od.period[0]);
                od.period[1] = fmin(thisPeriod, od.period[1]);
                runningMean(&od.period[2], thisPeriod, od.eventcount - 1);
            }
            od.tLastEvent = tThisEvent;
            od.thisNMaxima = 0;
            return od.eventcount >= op.maxEventCount;
        } else {  // nhood2 :178-205
            ++od.eventcount;
            R tThisEvent = ti;
            if (od.eventcount > 1) {
                od.nMaxima[0] = fmax((R)od.thisNMaxima, od.nMaxima[0]);
                od.nMaxima[1] = fmin((R)od.thisNMaxima, od.nMaxima[1]);
                runningMean(&od.nMaxima[2], (R)od.thisNMaxima, od.eventcount - 1);
                R thisPeriod = tThisEvent - od.tLastEvent;
                od.period[0] = fmax(thisPeriod, od.period[0]);
                od.period[1] = fmin(thisPeriod, od.period[1]);
                runningMean(&od.period[2], thisPeriod, od.eventcount - 1);
            }
            od.tLastEvent = tThisEvent;
            od.thisNMaxima = 0;
            if (od.eventcount <= (unsigned)S)
                od.list[0][od.eventcount - 1] = tThisEvent;
            return od.eventcount == op.maxEventCount;
        }
    }

    // finalizeFeatures: feature vector in the reference's order
    void finalizeFeatures(const ObserverData<R> &od, std::vector<R> &F) const
    {
        F.clear();
        bool p = od.eventcount > 1;
        auto put3 = [&](const R *s) {
            for (int k = 0; k < 3; ++k)
                F.push_back(p ? s[k] : R(0));
        };
        if (kind == OB_THRESH2) {  // :411-456
            put3(od.period);
            put3(od.nMaxima);
            put3(od.upDuration);
            put3(od.downDuration);
            put3(od.duty);
            put3(od.activeDip);
            for (int j = 0; j < N; ++j) {
                F.push_back(od.xTrajectoryMax[j]);
                F.push_back(od.xTrajectoryMin[j]);
                F.push_back(od.xTrajectoryMean[j]);
                F.push_back(od.dxTrajectoryMax[j]);
                F.push_back(od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                F.push_back(od.auxTrajectoryMax[j]);
                F.push_back(od.auxTrajectoryMin[j]);
                F.push_back(od.auxTrajectoryMean[j]);
            }
            for (int j = 0; j < S; ++j) {
                F.push_back(od.list[0][j]);
                F.push_back(od.list[1][j]);
            }
            F.push_back((R)od.eventcount);
            F.push_back((R)od.stepcount);
            F.push_back(od.stepDt[0]);
            F.push_back(od.stepDt[1]);
            F.push_back(od.stepDt[2]);
        } else if (kind == OB_LOCALMAX) {  // :288-316
            put3(od.IMI);
            put3(od.amp);
            for (int j = 0; j < N; ++j) {
                F.push_back(od.xTrajectoryMax[j]);
                F.push_back(od.xTrajectoryMin[j]);
                F.push_back(od.xTrajectoryMean[j]);
                F.push_back(od.dxTrajectoryMax[j]);
                F.push_back(od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                F.push_back(od.auxTrajectoryMax[j]);
                F.push_back(od.auxTrajectoryMin[j]);
                F.push_back(od.auxTrajectoryMean[j]);
            }
            for (int j = 0; j < S; ++j)
                for (int L = 0; L < 4; ++L)
                    F.push_back(od.list[L][j]);
            F.push_back((R)od.eventcount);
            F.push_back((R)od.stepcount);
        } else if (kind == OB_BASIC) {  // observer_basic.clh:77-85
            const unsigned f = op.fVarIx;
            F.push_back(od.xTrajectoryMax[f]);
            F.push_back(od.xTrajectoryMin[f]);
            F.push_back(od.xTrajectoryMean[f]);
            F.push_back(od.dxTrajectoryMax[f]);
            F.push_back(od.dxTrajectoryMin[f]);
            F.push_back((R)od.stepcount);
        } else if (kind == OB_BASICALL) {  // observer_basic_allVar.clh:114-130
            for (int j = 0; j < N; ++j) {
                F.push_back(od.xTrajectoryMax[j]);
                F.push_back(od.xTrajectoryMin[j]);
                F.push_back(od.xTrajectoryMean[j]);
                F.push_back(od.dxTrajectoryMax[j]);
                F.push_back(od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                F.push_back(od.auxTrajectoryMax[j]);
                F.push_back(od.auxTrajectoryMin[j]);
                F.push_back(od.auxTrajectoryMean[j]);
            }
            F.push_back((R)od.stepcount);
        } else if (kind == OB_NHOOD1) {  // observer_neighborhood_1.clh:279-306
            put3(od.period);
            put3(od.nMaxima);
            for (int j = 0; j < N; ++j) {
                F.push_back(od.xTrajectoryMax[j]);
                F.push_back(od.xTrajectoryMin[j]);
                F.push_back(od.xTrajectoryMean[j]);
                F.push_back(od.dxTrajectoryMax[j]);
                F.push_back(od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                F.push_back(od.auxTrajectoryMax[j]);
                F.push_back(od.auxTrajectoryMin[j]);
                F.push_back(od.auxTrajectoryMean[j]);
            }
            F.push_back((R)(unsigned int)(od.eventcount - 1u));  // periods; wraps at eventcount 0
            F.push_back((R)od.stepcount);
            F.push_back(od.stepDt[0]);
            F.push_back(od.stepDt[1]);
            F.push_back(od.stepDt[2]);
        } else {  // nhood2 :271-302
            put3(od.period);
            put3(od.nMaxima);
            for (int j = 0; j < N; ++j) {
                F.push_back(od.xTrajectoryMax[j]);
                F.push_back(od.xTrajectoryMin[j]);
                F.push_back(od.xTrajectoryMean[j]);
                F.push_back(od.xTrajectoryRange[j]);
                F.push_back(od.x0[j]);
                F.push_back(od.dxTrajectoryMax[j]);
                F.push_back(od.dxTrajectoryMin[j]);
            }
            for (int j = 0; j < A; ++j) {
                F.push_back(od.auxTrajectoryMax[j]);
                F.push_back(od.auxTrajectoryMin[j]);
                F.push_back(od.auxTrajectoryMean[j]);
            }
            for (int j = 0; j < S; ++j)
                F.push_back(od.list[0][j]);
            F.push_back((R)od.eventcount);
            F.push_back((R)od.stepcount);
            F.push_back(od.stepDt[0]);
            F.push_back(od.stepDt[1]);
            F.push_back(od.stepDt[2]);
        }
    }

    // finalizeObserverData: shift stored times by -(tf - t0)
    // (observer_threshold_2.clh:458-470, observer_local_maximum.clh:319-328,
    // observer_neighborhood_2.clh:305-313).  CLODE_ORACLE_MUTANT_NO_TIME_SHIFT
    // builds a deliberately wrong checker without the shift, which the
    // continuation tests must tell apart from this one.
    void finalizeObserverData(R ti, R t0, ObserverData<R> &od) const
    {
#ifdef CLODE_ORACLE_MUTANT_NO_TIME_SHIFT
        return;
#endif
        R T = ti - t0;
        od.t_start -= T;
        if (kind == OB_THRESH2) {
            od.tLastEvent -= T;
            od.tThisDown -= T;
            od.tLastMax -= T;
            od.tLastMin -= T;
        } else if (kind == OB_LOCALMAX) {
            od.tLastMax = od.tLastMax - T;
            od.tLastMin = od.tLastMin - T;
        } else if (kind == OB_NHOOD1 || kind == OB_NHOOD2) {
            od.tLastEvent -= T;
        }
        for (int j = 0; j < 3; ++j)
            od.tbuffer[j] -= T;
    }

    bool twoPass() const { return kind == OB_THRESH2 || kind == OB_NHOOD2; }

    int nFeatures() const
    {
        if (kind == OB_THRESH2)
            return 18 + 5 * N + 3 * A + 2 * S + 5;
        if (kind == OB_LOCALMAX)
            return 6 + 5 * N + 3 * A + 4 * S + 2;
        if (kind == OB_BASIC)
            return 6;
        if (kind == OB_BASICALL)
            return 5 * N + 3 * A + 1;
        if (kind == OB_NHOOD1)
            return 6 + 5 * N + 3 * A + 5;
        return 6 + 7 * N + 3 * A + S + 5;
    }
};

#undef OC

}  // namespace oracle
