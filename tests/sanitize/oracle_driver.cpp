// Drives the CPU oracle through every kind x stepper x observer x precision,
// with max_store 0, 1 and 7 and per-member spans; built with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/test_oracle_sanitized.py.
#include <cstdio>
#include <cstdint>
#include <vector>
#include <cmath>
extern "C" {
void *oracle_create(int, int, int, int, int, int, int, int, int);
void oracle_destroy(void *);
int oracle_set_threads(void *, int);
int oracle_set_problem_data(void *, const double *, size_t, const double *, size_t);
int oracle_set_tspan(void *, double, double);
int oracle_set_solver_params(void *, double, double, double, double, unsigned, unsigned, unsigned, int);
int oracle_set_observer_params(void *, const unsigned *, const double *);
int oracle_seed_rng(void *, int);
int oracle_transient(void *);
int oracle_features(void *, int);
int oracle_trajectory(void *);
int oracle_set_member_tspan(void *, const double *, size_t);
}
int main() {
    const int nv = NV, np = NP, na = NA, nw = NW;
    int fails = 0;
    for (int prec = 1; prec <= 2; ++prec)
    for (int st = 1; st <= 6; ++st) {
        if ((st == 6) != (nw > 0)) continue;
        for (int kind = 0; kind <= 2; ++kind)
        for (int ob = (kind == 1 ? 1 : 0); ob <= (kind == 1 ? 6 : 0); ++ob)
        for (int nstore : {0, 3}) {
            if (kind != 1 && nstore) continue;
            void *h = oracle_create(kind, prec, st, ob, nv, np, na, nw, nstore);
            oracle_set_threads(h, 2);
            const int n = 37;
            std::vector<double> x0(nv * n), p(np * n);
            for (int i = 0; i < nv * n; ++i) x0[i] = X0SCALE * (0.5 + 0.01 * (i % 17));
            for (int i = 0; i < np * n; ++i) p[i] = PSCALE * (0.2 + 0.05 * (i % 11));
            oracle_set_problem_data(h, x0.data(), x0.size(), p.data(), p.size());
            unsigned u4[4] = {0, 0, 100, (unsigned)nstore};
            double r8[8] = {0, 0, 0.05, 0.3, 0.2, 0, 0, 0};
            if (kind == 1) oracle_set_observer_params(h, u4, r8);
            for (unsigned ms : {0u, 1u, 7u}) {
                oracle_set_solver_params(h, 0.05, 0.5, 1e-6, 1e-3, 2000, ms, 1, 0);
                oracle_seed_rng(h, 5);
                oracle_set_tspan(h, 0.0, TEND);
                if (kind == 0) fails += oracle_transient(h) != 0;
                if (kind == 1) { fails += oracle_features(h, 1) != 0; fails += oracle_features(h, -1) != 0; }
                if (kind == 2) fails += oracle_trajectory(h) != 0;
            }
            std::vector<double> tm(2 * n);
            for (int i = 0; i < n; ++i) { tm[i] = 0.1 * i; tm[n + i] = tm[i] + 1.0 + 0.3 * i; }
            oracle_set_member_tspan(h, tm.data(), tm.size());
            if (kind == 0) fails += oracle_transient(h) != 0;
            if (kind == 1) fails += oracle_features(h, 1) != 0;
            if (kind == 2) fails += oracle_trajectory(h) != 0;
            oracle_destroy(h);
        }
    }
    std::printf("done, failures %d\n", fails);
    return 0;
}
