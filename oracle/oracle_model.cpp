// CPU oracle: one shared library per right-hand-side model.
//
// TEST INFRASTRUCTURE ONLY (see clode_oracle.hpp).  Built by
// oracle/build_oracle.py with -DCLODE_ORACLE_RHS="<model>.cl"; exposes a
// stateful ensemble with the reference's host semantics through extern "C"
// functions used by tests/oracle.py over ctypes.
//
// Host semantics restated from clode/cpp/CLODE.cpp and CLODEfeatures.cpp:
//  * the per-member step-size buffer is filled with sp.dt only when the
//    number of members changes (CLODE.cpp:184-202, H3), not by
//    setSolverParams (CLODE.cpp:377-403);
//  * seedRNG(int s) sets word k of the 2*nPts state words to s+k
//    (CLODE.cpp:447-464);
//  * observer data persist between features() calls until nPts changes or
//    features(true) is called (CLODEfeatures.cpp:143-259);
//  * shiftX0 copies xf to x0; shiftTspan moves (t0,t1) to (t1, 2 t1 - t0).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define CM_FN inline
#include "../clode_amd/csrc/device/cm_math.h"
#ifdef CLODE_ORACLE_COUNTING
#include "flop_count.hpp"
#endif
#include "clode_oracle.hpp"
#include "oracle_ensemble.hpp"
#ifdef CLODE_ORACLE_COUNTING
template <>
struct oracle::Consts<fc::Counted> {
    static fc::Counted big() { return fc::Counted(1.7976931348623157e308); }
    static fc::Counted c(double d, float) { return fc::Counted(d); }
};
#endif

// ---- OpenCL C built-ins for the right-hand side, host versions -----------
static inline double sinpi(double x) { return std::sin(3.141592653589793238462643383279502884 * x); }
static inline float sinpi(float x) { return std::sin(3.14159265358979323846f * x); }
static inline double cospi(double x) { return std::cos(3.141592653589793238462643383279502884 * x); }
static inline float cospi(float x) { return std::cos(3.14159265358979323846f * x); }
static inline double rsqrt(double x) { return 1.0 / std::sqrt(x); }
static inline float rsqrt(float x) { return 1.0f / std::sqrt(x); }
using std::acos;
using std::acosh;
using std::asin;
using std::asinh;
using std::atan;
using std::atan2;
using std::atanh;
using std::cbrt;
using std::ceil;
using std::copysign;
using std::cos;
using std::cosh;
using std::erf;
using std::erfc;
using std::exp2;
using std::expm1;
using std::fabs;
using std::fdim;
using std::floor;
using std::fmax;
using std::fmin;
using std::fmod;
using std::hypot;
using std::ilogb;
using std::ldexp;
using std::lgamma;
using std::log10;
using std::log1p;
using std::log2;
using std::nextafter;
using std::remainder;
using std::rint;
using std::round;
using std::sin;
using std::sinh;
using std::sqrt;
using std::tan;
using std::tanh;
using std::tgamma;
using std::trunc;
static inline float exp10(float x) { return ::exp10f(x); }

#include "../clode_amd/csrc/device/clode_builtins.h"

#ifndef CLODE_ORACLE_RHS
#error "CLODE_ORACLE_RHS must name the right-hand-side source"
#endif

namespace rhs32 {
typedef float realtype;
#define RCONST(x) (x##f)
#define ZERO 0.0f
#define ONE 1.0f
#define BIG_REAL 3.40282347e+38f
#include CLODE_ORACLE_RHS
#undef RCONST
#undef ZERO
#undef ONE
#undef BIG_REAL
}  // namespace rhs32

namespace rhs64 {
typedef double realtype;
#define RCONST(x) (x)
#define ZERO 0.0
#define ONE 1.0
#define BIG_REAL 1.7976931348623157e308
#include CLODE_ORACLE_RHS
#undef RCONST
#undef ZERO
#undef ONE
#undef BIG_REAL
}  // namespace rhs64

#ifdef CLODE_ORACLE_COUNTING
// the right-hand side on the operation-counting type (flop_count.hpp)
namespace rhsC {
typedef fc::Counted realtype;
#define RCONST(x) (x)
#define ZERO 0.0
#define ONE 1.0
#define BIG_REAL 1.7976931348623157e308
#include CLODE_ORACLE_RHS
#undef RCONST
#undef ZERO
#undef ONE
#undef BIG_REAL
}  // namespace rhsC
#endif

#undef exp
#undef log
#undef pow
#undef pown

using namespace oracle;

template <>
struct oracle_drv::RhsOf<float> {
    static void call(const float t, const float *x, const float *p, float *dx, float *aux, const float *w)
    {
        rhs32::getRHS(t, x, p, dx, aux, w);
    }
};
template <>
struct oracle_drv::RhsOf<double> {
    static void call(const double t, const double *x, const double *p, double *dx, double *aux, const double *w)
    {
        rhs64::getRHS(t, x, p, dx, aux, w);
    }
};

#ifdef CLODE_ORACLE_COUNTING
template <>
struct oracle_drv::RhsOf<fc::Counted> {
    static void call(const fc::Counted t, const fc::Counted *x, const fc::Counted *p, fc::Counted *dx,
                     fc::Counted *aux, const fc::Counted *w)
    {
        rhsC::getRHS(t, x, p, dx, aux, w);
    }
};
#endif

namespace {

using oracle_drv::Config;
using oracle_drv::Ensemble;

struct Handle {
    Config cfg;
    Ensemble<float> *f32 = nullptr;
    Ensemble<double> *f64 = nullptr;
#ifdef CLODE_ORACLE_COUNTING
    Ensemble<fc::Counted> *cnt = nullptr;  // precision 9: operation counting
#endif
};

template <class Fn>
int with(void *h, Fn fn)
{
    Handle *H = (Handle *)h;
    if (!H)
        return 1;
#ifdef CLODE_ORACLE_COUNTING
    if (H->cnt)
        return fn(*H->cnt);
#endif
    if (H->f32)
        return fn(*H->f32);
    return fn(*H->f64);
}

template <class R>
std::vector<double> todouble(const std::vector<R> &v)
{
    return std::vector<double>(v.begin(), v.end());
}

}  // namespace

extern "C" {

void *oracle_create(int kind, int prec, int stepper, int observer, int nVar, int nPar, int nAux, int nWiener,
                    int nStore)
{
    if (nVar <= 0 || nVar > oracle::MAXV)
        return nullptr;
    Handle *H = new Handle;
    H->cfg = Config{kind, prec, stepper, observer, nVar, nPar, nAux, nWiener, nStore, 1};
#ifdef CLODE_ORACLE_COUNTING
    if (prec == 9) {
        H->cnt = new Ensemble<fc::Counted>(H->cfg);
        fc::reset();
        return H;
    }
#endif
    if (prec == 1)
        H->f32 = new Ensemble<float>(H->cfg);
    else
        H->f64 = new Ensemble<double>(H->cfg);
    return H;
}

void oracle_destroy(void *h)
{
    Handle *H = (Handle *)h;
    if (!H)
        return;
    delete H->f32;
    delete H->f64;
#ifdef CLODE_ORACLE_COUNTING
    delete H->cnt;
#endif
    delete H;
}

#ifdef CLODE_ORACLE_COUNTING
// Operation counts since the counting ensemble was created, in the layout
// [phase 4][bucket 4][count 10] then [phase 4][trials 3]: phases transient,
// warm-up, features, trajectory; buckets rejected, accepted, failed trials
// and the whole phase; counts add, mul, div, fma, exp, log, pow, pown, sqrt,
// other transcendental; trials rejected, accepted, failed.
int oracle_counts(double *out, size_t n)
{
    const size_t need = (size_t)fc::N_PHASE * fc::N_BUCKET * fc::N_COUNT + (size_t)fc::N_PHASE * 3;
    if (!out || n < need)
        return 1;
    fc::close_phase();
    size_t k = 0;
    for (int p = 0; p < fc::N_PHASE; ++p)
        for (int b = 0; b < fc::N_BUCKET; ++b)
            for (int c = 0; c < fc::N_COUNT; ++c)
                out[k++] = (double)fc::st().bucket[p][b].c[c];
    for (int p = 0; p < fc::N_PHASE; ++p)
        for (int t = 0; t < 3; ++t)
            out[k++] = (double)fc::st().trials[p][t];
    return 0;
}
#endif

int oracle_set_threads(void *h, int n)
{
    return with(h, [&](auto &e) {
#ifdef CLODE_ORACLE_COUNTING
        e.cfg.nthreads = 1;  // one running tally
        (void)n;
#else
        e.cfg.nthreads = n;
#endif
        return 0;
    });
}

int oracle_set_problem_data(void *h, const double *x0, size_t nx0, const double *pars, size_t np)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        int n = (int)(nx0 / e.cfg.nVar);
        if (nx0 % e.cfg.nVar || (e.cfg.nPar > 0 && np != (size_t)n * e.cfg.nPar))
            return 1;
        e.setNpts(n);
        for (size_t k = 0; k < nx0; ++k)
            e.x0[k] = (R)x0[k];
        for (size_t k = 0; k < np; ++k)
            e.pars[k] = (R)pars[k];
        return 0;
    });
}

int oracle_set_x0(void *h, const double *x0, size_t n)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        if (n != e.x0.size())
            return 1;
        for (size_t k = 0; k < n; ++k)
            e.x0[k] = (R)x0[k];
        return 0;
    });
}

int oracle_set_tspan(void *h, double t0, double t1)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        e.tspan_d[0] = t0;
        e.tspan_d[1] = t1;
        e.t0 = (R)t0;
        e.t1 = (R)t1;
        return 0;
    });
}

int oracle_set_solver_params(void *h, double dt, double dtmax, double abstol, double reltol, unsigned max_steps,
                             unsigned max_store, unsigned nout, int refresh_dt)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        e.sp_dt_d = dt;
        e.sp.dt = (R)dt;
        e.sp.dtmax = (R)dtmax;
        e.sp.abstol = (R)abstol;
        e.sp.reltol = (R)reltol;
        e.sp.max_steps = max_steps;
        e.sp.max_store = max_store;
        e.sp.nout = nout;
        if (refresh_dt)
            std::fill(e.dt.begin(), e.dt.end(), (R)dt);
        return 0;
    });
}

int oracle_set_observer_params(void *h, const unsigned *u4, const double *r8)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        e.op.eVarIx = u4[0];
        e.op.fVarIx = u4[1];
        e.op.maxEventCount = u4[2];
        e.op.maxEventTimestamps = u4[3];
        e.op.minXamp = (R)r8[0];
        e.op.minIMI = (R)r8[1];
        e.op.nHoodRadius = (R)r8[2];
        e.op.xUpThresh = (R)r8[3];
        e.op.xDownThresh = (R)r8[4];
        e.op.dxUpThresh = (R)r8[5];
        e.op.dxDownThresh = (R)r8[6];
        e.op.eps_dx = (R)r8[7];
        return 0;
    });
}

int oracle_seed_rng(void *h, int seed)
{
    return with(h, [&](auto &e) {
        for (size_t k = 0; k < e.rng.size(); ++k)
            e.rng[k] = (uint64_t)(int64_t)(int32_t)((uint32_t)seed + (uint32_t)k);
        return 0;
    });
}

int oracle_set_rng(void *h, const uint64_t *w, size_t n)
{
    return with(h, [&](auto &e) {
        if (n != e.rng.size())
            return 1;
        std::copy(w, w + n, e.rng.begin());
        return 0;
    });
}

int oracle_set_dt(void *h, const double *d, size_t n)
{
    return with(h, [&](auto &e) {
        using R = typename std::decay_t<decltype(e.x0)>::value_type;
        if (n != e.dt.size())
            return 1;
        for (size_t k = 0; k < n; ++k)
            e.dt[k] = (R)d[k];
        return 0;
    });
}

int oracle_transient(void *h)
{
    return with(h, [&](auto &e) {
        CLODE_ORACLE_PHASE(0);
        e.forEach([&e](int i) { return e.transientMember(i); });
        return 0;
    });
}

int oracle_initialize_observer(void *h)
{
    return with(h, [&](auto &e) {
        e.initializeObserver();
        return 0;
    });
}

int oracle_features(void *h, int reinit)
{
    return with(h, [&](auto &e) {
        e.features(reinit);
        return 0;
    });
}

int oracle_trajectory(void *h)
{
    return with(h, [&](auto &e) {
        e.trajectory();
        return 0;
    });
}

// per-member time spans [2][nPts] (row 0 start, row 1 end); n == 0 clears
int oracle_set_member_tspan(void *h, const double *t, size_t n)
{
    return with(h, [&](auto &e) {
        if (n == 0) {
            e.tspanM.clear();
            return 0;
        }
        if (n != 2 * (size_t)e.nPts)
            return 1;
        e.tspanM.assign(t, t + n);
        return 0;
    });
}

int oracle_shift_x0(void *h)
{
    return with(h, [&](auto &e) {
        e.x0 = e.xf;
        return 0;
    });
}

unsigned long long oracle_last_steps(void *h)
{
    unsigned long long s = 0;
    with(h, [&](auto &e) {
        s = e.lastSteps;
        return 0;
    });
    return s;
}

unsigned long long oracle_last_trials(void *h)
{
    unsigned long long s = 0;
    with(h, [&](auto &e) {
        s = e.lastTrials;
        return 0;
    });
    return s;
}

unsigned long long oracle_last_warm_steps(void *h)
{
    unsigned long long s = 0;
    with(h, [&](auto &e) {
        s = e.warmSteps;
        return 0;
    });
    return s;
}

// name: x0 xf pars dt tf F t x dx aux
long long oracle_size(void *h, const char *name)
{
    long long n = -1;
    with(h, [&](auto &e) {
        std::string s(name);
        if (s == "x0") n = e.x0.size();
        else if (s == "xf") n = e.xf.size();
        else if (s == "pars") n = e.pars.size();
        else if (s == "dt") n = e.dt.size();
        else if (s == "tf") n = e.tf.size();
        else if (s == "F") n = e.F.size();
        else if (s == "t") n = e.tt.size();
        else if (s == "x") n = e.tx.size();
        else if (s == "dx") n = e.tdx.size();
        else if (s == "aux") n = e.taux.size();
        else if (s == "nStored") n = e.nStored.size();
        else if (s == "rng") n = e.rng.size();
        return 0;
    });
    return n;
}

int oracle_get(void *h, const char *name, double *out, size_t n)
{
    return with(h, [&](auto &e) {
        std::string s(name);
        std::vector<double> v;
        if (s == "x0") v = todouble(e.x0);
        else if (s == "xf") v = todouble(e.xf);
        else if (s == "pars") v = todouble(e.pars);
        else if (s == "dt") v = todouble(e.dt);
        else if (s == "tf") v = todouble(e.tf);
        else if (s == "F") v = todouble(e.F);
        else if (s == "t") v = todouble(e.tt);
        else if (s == "x") v = todouble(e.tx);
        else if (s == "dx") v = todouble(e.tdx);
        else if (s == "aux") v = todouble(e.taux);
        else if (s == "nStored") v = std::vector<double>(e.nStored.begin(), e.nStored.end());
        else return 1;
        if (n < v.size())
            return 2;
        std::copy(v.begin(), v.end(), out);
        return 0;
    });
}

int oracle_get_rng(void *h, uint64_t *out, size_t n)
{
    return with(h, [&](auto &e) {
        if (n < e.rng.size())
            return 2;
        std::copy(e.rng.begin(), e.rng.end(), out);
        return 0;
    });
}

// single-call RNG helpers for known-answer tests
void oracle_xoroshiro_next(uint64_t *s, uint64_t *out, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        out[k] = oracle::next(s);
}

void oracle_randn(int prec, uint64_t *s, double *out, size_t n)
{
    if (prec == 1) {
        RngData<float> rd{{s[0], s[1]}, false, 0.f};
        for (size_t k = 0; k < n; ++k)
            out[k] = randn(rd);
        s[0] = rd.state[0];
        s[1] = rd.state[1];
    } else {
        RngData<double> rd{{s[0], s[1]}, false, 0.0};
        for (size_t k = 0; k < n; ++k)
            out[k] = randn(rd);
        s[0] = rd.state[0];
        s[1] = rd.state[1];
    }
}

// elementary functions, for accuracy tests against glibc
void oracle_cm_eval(int fn, int prec, const double *x, const double *y, double *out, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        if (prec == 1) {
            float a = (float)x[k], b = y ? (float)y[k] : 0.f;
            out[k] = fn == 0 ? cm::exp(a) : fn == 1 ? cm::log(a) : fn == 2 ? cm::pow(a, b) : fn == 3 ? cm::pown(a, (int)b)
                   : fn == 4 ? cm::root<3>(a) : fn == 5 ? cm::root<5>(a) : fn == 7 ? cm::root<2>(a)
                   : cm::min_step(a, b);
        } else {
            double a = x[k], b = y ? y[k] : 0.0;
            out[k] = fn == 0 ? cm::exp(a) : fn == 1 ? cm::log(a) : fn == 2 ? cm::pow(a, b) : fn == 3 ? cm::pown(a, (int)b)
                   : fn == 4 ? cm::root<3>(a) : fn == 5 ? cm::root<5>(a) : fn == 7 ? cm::root<2>(a)
                   : cm::min_step(a, b);
        }
    }
}

}  // extern "C"
