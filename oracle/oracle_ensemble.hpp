// Drivers of the CPU oracle: one ensemble of one model, the reference's
// host semantics and kernels member by member (transient, observer warm-up,
// features, trajectory).  TEST INFRASTRUCTURE ONLY, shared by the oracle
// library (oracle_model.cpp) and the operation-counting build
// (flop_count.cpp); the including file supplies RhsOf<R>::call for its real
// types and must include clode_oracle.hpp first.
//   transient.cl:9-77, initializeObserver.cl:9-84, features.cl:10-106,
//   trajectory.cl:14-112 (clODE repository, clode/cpp/)
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

// Marks the start of a pass (0 transient, 1 observer warm-up, 2 features,
// 3 trajectory); empty except in the operation-counting build.
#ifndef CLODE_ORACLE_PHASE
#define CLODE_ORACLE_PHASE(p)
#endif

namespace oracle_drv {

using namespace oracle;
using std::sqrt;

template <class R>
struct RhsOf;

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
    unsigned long long lastTrials = 0; // step trials of the last pass (steps plus rejected trials)
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
        mark_varying(m.ti);
        if (cfg.stepper == ST_BS23 || cfg.stepper == ST_DOPRI5 || cfg.stepper == ST_HE12)
            mark_varying(m.dt);  // a fixed step size is loop-invariant
        for (int j = 0; j < cfg.nPar; ++j)
            m.p[j] = pars[(size_t)j * nPts + i];
        for (int j = 0; j < N; ++j) {
            m.x[j] = x0[(size_t)j * nPts + i];
            mark_varying(m.x[j]);
        }
        m.rd.state[0] = rng[i];
        m.rd.state[1] = rng[(size_t)nPts + i];
        m.rd.randnUselast = false;
        m.rd.randnLast = 0;
        for (int j = 0; j < cfg.nWiener; ++j)
            m.w[j] = cfg.stepper == ST_SEULER ? randn(m.rd) / sqrt(m.dt) : R(0);
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
            mark_varying(m.ti);
            for (int j = 0; j < cfg.nVar; ++j) {
                m.x[j] = x0[(size_t)j * nPts + i];
                mark_varying(m.x[j]);
            }
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
        unsigned long long total = 0, trials = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(cfg.nthreads > 0 ? cfg.nthreads : 1) reduction(+ : total, trials)
        for (int i = 0; i < nPts; ++i) {
            const unsigned long long t0 = trial_count();
            total += fn(i);
            trials += trial_count() - t0;
        }
        lastSteps = total;
        lastTrials = trials;
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
        CLODE_ORACLE_PHASE(1);
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
        CLODE_ORACLE_PHASE(2);
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
        CLODE_ORACLE_PHASE(3);
        forEach([this](int i) { return trajectoryMember(i); });
    }
};

}  // namespace oracle_drv
