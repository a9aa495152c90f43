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
#include "clode_oracle.hpp"

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

#undef exp
#undef log
#undef pow
#undef pown

using namespace oracle;

namespace {

template <class R>
struct RhsOf;
template <>
struct RhsOf<float> {
    static void call(const float t, const float *x, const float *p, float *dx, float *aux, const float *w)
    {
        rhs32::getRHS(t, x, p, dx, aux, w);
    }
};
template <>
struct RhsOf<double> {
    static void call(const double t, const double *x, const double *p, double *dx, double *aux, const double *w)
    {
        rhs64::getRHS(t, x, p, dx, aux, w);
    }
};

struct Config {
    int kind, prec, stepper, observer, nVar, nPar, nAux, nWiener, nStore, nthreads;
};

template <class R>
struct Ensemble {
    Config cfg;
    Problem<R> pb;
    int nPts = 0;
    std::vector<R> x0, xf, pars, dt, tf;
    std::vector<uint64_t> rng;
    R t0 = 0, t1 = 1000;
    double tspan_d[2] = {0, 1000};
    std::vector<double> tspanM;  // optional per-member spans [2][nPts]: row 0 start, row 1 end
    R mt0(int i) const { return tspanM.empty() ? t0 : (R)tspanM[i]; }
    R mt1(int i) const { return tspanM.empty() ? t1 : (R)tspanM[(size_t)nPts + i]; }
    SolverParamsT<R> sp{};
    double sp_dt_d = 0.1;
    ObserverParamsT<R> op{};
    std::vector<ObserverData<R>> od;
    bool obsInit = false;
    std::vector<R> F;
    std::vector<R> tt, tx, tdx, taux;
    std::vector<int> nStored;
    unsigned int storeAlloc = 0;
    unsigned long long lastSteps = 0;  // member-steps of the last pass
    unsigned long long warmSteps = 0;  // member-steps of the observer warm-up run by the last features()

    explicit Ensemble(const Config &c) : cfg(c)
    {
        pb.nVar = c.nVar;
        pb.nPar = c.nPar;
        pb.nAux = c.nAux;
        pb.nWiener = c.nWiener;
        pb.rhs = &RhsOf<R>::call;
        sp.dt = (R)0.1;
        sp.dtmax = (R)0.5;
        sp.abstol = (R)1e-6;
        sp.reltol = (R)1e-3;
        sp.max_steps = 1000000;
        sp.max_store = 1000000;
        sp.nout = 1;
    }

    Observer<R> observer() const
    {
        Observer<R> o;
        o.kind = cfg.observer;
        o.N = cfg.nVar;
        o.A = cfg.nAux;
        o.S = cfg.nStore;
        o.op = op;
        return o;
    }

    void setNpts(int n)
    {
        if (n == nPts)
            return;
        nPts = n;
        x0.assign((size_t)cfg.nVar * n, 0);
        xf.assign((size_t)cfg.nVar * n, 0);
        pars.assign((size_t)cfg.nPar * n, 0);
        dt.assign(n, (R)sp_dt_d);
        tf.assign(n, 0);
        rng.assign(2 * (size_t)n, 0);
        od.clear();
        F.clear();
        obsInit = false;
        storeAlloc = 0;
        tspanM.clear();
    }

    struct Member {
        std::vector<R> p, x, dx, aux, w;
        R ti, dt;
        RngData<R> rd;
    };

    void load(Member &m, int i) const
    {
        const int N = cfg.nVar;
        m.p.assign(std::max(1, cfg.nPar), 0);
        m.x.assign(N, 0);
        m.dx.assign(N, 0);
        m.aux.assign(std::max(1, cfg.nAux), 0);
        m.w.assign(std::max(1, cfg.nWiener), 0);
        m.ti = mt0(i);
        m.dt = dt[i];
        for (int j = 0; j < cfg.nPar; ++j)
            m.p[j] = pars[(size_t)j * nPts + i];
        for (int j = 0; j < N; ++j)
            m.x[j] = x0[(size_t)j * nPts + i];
        m.rd.state[0] = rng[i];
        m.rd.state[1] = rng[(size_t)nPts + i];
        m.rd.randnUselast = false;
        m.rd.randnLast = 0;
        for (int j = 0; j < cfg.nWiener; ++j)
            m.w[j] = cfg.stepper == ST_SEULER ? randn(m.rd) / std::sqrt(m.dt) : R(0);
    }

    void store(const Member &m, int i)
    {
        for (int j = 0; j < cfg.nVar; ++j)
            xf[(size_t)j * nPts + i] = m.x[j];
        rng[i] = m.rd.state[0];
        rng[(size_t)nPts + i] = m.rd.state[1];
        dt[i] = m.dt;
        tf[i] = m.ti;
    }

    // transient.cl:9-77
    unsigned int transientMember(int i)
    {
        Integrator<R> ig(pb, cfg.stepper);
        Member m;
        load(m, i);
        pb.rhs(m.ti, m.x.data(), m.p.data(), m.dx.data(), m.aux.data(), m.w.data());
        const R t1 = mt1(i);
        unsigned int step = 0;
        while (m.ti <= t1 && step < sp.max_steps) {
            ++step;
            ig.step(&m.ti, m.x.data(), m.dx.data(), m.p.data(), sp, &m.dt, t1, m.aux.data(), m.w.data(), m.rd);
        }
        store(m, i);
        return step;
    }

    // initializeObserver.cl:9-84
    unsigned int initObserverMember(int i)
    {
        Integrator<R> ig(pb, cfg.stepper);
        Observer<R> ob = observer();
        Member m;
        load(m, i);
        pb.rhs(m.ti, m.x.data(), m.p.data(), m.dx.data(), m.aux.data(), m.w.data());
        ObserverData<R> &d = od[i];
        ob.init(m.ti, m.x.data(), m.dx.data(), d);
        const R t0 = mt0(i), t1 = mt1(i);
        unsigned int step = 0;
        if (ob.twoPass()) {
            while (m.ti < t1 && step < sp.max_steps) {
                ++step;
                ig.step(&m.ti, m.x.data(), m.dx.data(), m.p.data(), sp, &m.dt, t1, m.aux.data(), m.w.data(), m.rd);
                ob.warmup(m.ti, m.x.data(), m.dx.data(), d);
            }
            m.ti = t0;
            for (int j = 0; j < cfg.nVar; ++j)
                m.x[j] = x0[(size_t)j * nPts + i];
            pb.rhs(m.ti, m.x.data(), m.p.data(), m.dx.data(), m.aux.data(), m.w.data());
        }
        ob.initEventDetector(m.ti, m.x.data(), m.dx.data(), d);
        return step;
    }

    // features.cl:10-106
    unsigned int featuresMember(int i)
    {
        Integrator<R> ig(pb, cfg.stepper);
        Observer<R> ob = observer();
        Member m;
        load(m, i);
        ObserverData<R> &d = od[i];
        pb.rhs(m.ti, m.x.data(), m.p.data(), m.dx.data(), m.aux.data(), m.w.data());
        const R t0 = mt0(i), t1 = mt1(i);
        unsigned int step = 0;
        while (m.ti <= t1 && step < sp.max_steps) {
            ++step;
            ig.step(&m.ti, m.x.data(), m.dx.data(), m.p.data(), sp, &m.dt, t1, m.aux.data(), m.w.data(), m.rd);
            ob.update(m.ti, m.x.data(), m.dx.data(), m.aux.data(), d);
            if (ob.eventFunction(m.ti, m.x.data(), m.dx.data(), d))
                if (ob.computeEventFeatures(m.ti, m.x.data(), m.dx.data(), d))
                    break;
        }
        std::vector<R> f;
        ob.finalizeFeatures(d, f);
        for (size_t k = 0; k < f.size(); ++k)
            F[k * nPts + i] = f[k];
        ob.finalizeObserverData(m.ti, t0, d);
        store(m, i);
        return step;
    }

    // trajectory.cl:14-112 (never stores index max_store)
    unsigned int trajectoryMember(int i)
    {
        Integrator<R> ig(pb, cfg.stepper);
        Member m;
        load(m, i);
        pb.rhs(m.ti, m.x.data(), m.p.data(), m.dx.data(), m.aux.data(), m.w.data());
        const size_t n = nPts;
        const int N = cfg.nVar, A = cfg.nAux;
        auto put = [&](unsigned int s) {
            tt[s * n + i] = m.ti;
            for (int j = 0; j < N; ++j) {
                tx[(s * N + j) * n + i] = m.x[j];
                tdx[(s * N + j) * n + i] = m.dx[j];
            }
            for (int j = 0; j < A; ++j)
                taux[(s * A + j) * n + i] = m.aux[j];
        };
        const R t1 = mt1(i);
        unsigned int storeix = 0;
        if (sp.max_store > 0)  // max_store 0: no sample buffer at all
            put(storeix);
        unsigned int step = 0;
        while (m.ti <= t1 && step < sp.max_steps && storeix < sp.max_store) {
            ++step;
            ig.step(&m.ti, m.x.data(), m.dx.data(), m.p.data(), sp, &m.dt, t1, m.aux.data(), m.w.data(), m.rd);
            if (step % sp.nout == 0) {
                ++storeix;
                if (storeix < sp.max_store)
                    put(storeix);
            }
        }
        nStored[i] = (int)storeix;
        store(m, i);
        return step;
    }

    template <class Fn>
    void forEach(Fn fn)
    {
        unsigned long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(cfg.nthreads > 0 ? cfg.nthreads : 1) reduction(+ : total)
        for (int i = 0; i < nPts; ++i)
            total += fn(i);
        lastSteps = total;
    }

    void ensureObserver()
    {
        Observer<R> ob = observer();
        size_t fe = (size_t)ob.nFeatures() * nPts;
        if (F.size() != fe || od.size() != (size_t)nPts) {
            F.assign(fe, 0);
            od.assign(nPts, ObserverData<R>());
            for (auto &d : od)
                d.resize(cfg.nVar, cfg.nAux, cfg.nStore);
            obsInit = false;
        }
    }

    void initializeObserver()
    {
        ensureObserver();
        forEach([this](int i) { return initObserverMember(i); });
        obsInit = true;
    }

    void features(int reinit)
    {
        if (reinit == 1)
            obsInit = false;
        else if (reinit == 0)
            obsInit = true;
        ensureObserver();
        warmSteps = 0;
        if (!obsInit) {
            initializeObserver();
            warmSteps = lastSteps;
        }
        forEach([this](int i) { return featuresMember(i); });
    }

    void trajectory()
    {
        size_t te = (size_t)sp.max_store * nPts;
        tt.assign(te, 0);
        tx.assign(te * cfg.nVar, 0);
        tdx.assign(te * cfg.nVar, 0);
        taux.assign(cfg.nAux > 0 ? te * cfg.nAux : 1, 0);
        nStored.assign(nPts, 0);
        forEach([this](int i) { return trajectoryMember(i); });
    }
};

struct Handle {
    Config cfg;
    Ensemble<float> *f32 = nullptr;
    Ensemble<double> *f64 = nullptr;
};

template <class Fn>
int with(void *h, Fn fn)
{
    Handle *H = (Handle *)h;
    if (!H)
        return 1;
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
    delete H;
}

int oracle_set_threads(void *h, int n)
{
    return with(h, [&](auto &e) {
        e.cfg.nthreads = n;
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
            out[k] = fn == 0 ? cm::exp(a) : fn == 1 ? cm::log(a) : fn == 2 ? cm::pow(a, b) : cm::pown(a, (int)b);
        } else {
            double a = x[k], b = y ? y[k] : 0.0;
            out[k] = fn == 0 ? cm::exp(a) : fn == 1 ? cm::log(a) : fn == 2 ? cm::pow(a, b) : cm::pown(a, (int)b);
        }
    }
}

}  // extern "C"
