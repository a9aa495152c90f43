// Operation counting for the CPU oracle.  TEST / MEASUREMENT INFRASTRUCTURE:
// compiled only into the counting build of oracle_model.cpp
// (CLODE_ORACLE_COUNTING, oracle/_build/flopcount_<model>.so), which
// oracle/gen_perfmodel.py runs to generate the operation table of
// clode_amd/perfmodel.py and tests/test_perfmodel.py runs to check it.
//
// The oracle's steppers, controller, observers and drivers and the model's
// right-hand side are instantiated with the real type `Counted`: a double
// that also records whether it varies during the integration.  Values loaded
// as state, time or step size, and every RNG word, vary; literals, tableau
// coefficients, solver, observer and model parameters and span ends do not.
// Counting rule (algorithmic FLOPs):
//   * +, -, *, / count 1 and fma counts 2, when at least one operand varies;
//     an operation on invariant operands only is folded or hoisted out of
//     the time loop by a compiler and counts 0;
//   * negation, fabs, fmin, fmax, nextafter, comparisons and integer work
//     count 0;
//   * exp, log, pow, pown and sqrt are counted separately, not as FLOPs.
// So a*b + c counts 2 whether or not it is contracted into an fma.
//
// Attribution: every counted operation is charged to the current phase
// (transient, warm-up, features, trajectory).  Within a phase the stepper's
// trials are cut out by the oracle's trial hooks (the step set-up and one
// trial: from the stepper's entry, or from the end of the previous rejected
// trial, to the trial's end), by outcome: accepted, rejected, failed at
// hmin.  What remains in the phase (observer updates and events, the initial
// right-hand side of each member, finalisation) is "other".
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace fc {

enum { N_COUNT = 10 };
enum { C_ADD, C_MUL, C_DIV, C_FMA, C_EXP, C_LOG, C_POW, C_POWN, C_SQRT, C_OTHER };

struct Tally {
    unsigned long long c[N_COUNT];
    void clear() { std::memset(c, 0, sizeof(c)); }
};

enum { PH_TRANSIENT, PH_WARMUP, PH_FEATURES, PH_TRAJECTORY, N_PHASE };
// bucket = the trial hook's outcome code (0 rejected, 1 accepted, 2 failed), then the phase total
enum { B_REJECTED, B_ACCEPTED, B_FAILED, B_ALL, N_BUCKET };

struct State {
    int phase = PH_TRANSIENT;
    Tally now;                           // running total of the current phase
    Tally mark;                          // `now` at the current trial's start
    Tally bucket[N_PHASE][N_BUCKET];     // trials by outcome, and the phase total
    unsigned long long trials[N_PHASE][3];
};

inline State &st()
{
    static State s;
    return s;
}

inline void count(int k, bool var)
{
    if (var)
        ++st().now.c[k];
}

struct Counted {
    double v = 0.0;
    bool var = false;
    Counted() = default;
    Counted(double d) : v(d) {}
    Counted(float d) : v(d) {}
    Counted(int d) : v(d) {}
    Counted(unsigned int d) : v(d) {}
    Counted(long d) : v((double)d) {}
    Counted(long long d) : v((double)d) {}
    Counted(unsigned long d, bool rngword = true) : v((double)d), var(rngword) {}  // RNG words (uint64_t)
    Counted(unsigned long long d) : v((double)d) {}
    static Counted make(double v, bool var)
    {
        Counted c(v);
        c.var = var;
        return c;
    }
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    explicit operator unsigned int() const { return (unsigned int)v; }
    Counted operator-() const { return make(-v, var); }
    Counted operator+() const { return *this; }
};

inline Counted operator+(Counted a, Counted b) { count(C_ADD, a.var || b.var); return Counted::make(a.v + b.v, a.var || b.var); }
inline Counted operator-(Counted a, Counted b) { count(C_ADD, a.var || b.var); return Counted::make(a.v - b.v, a.var || b.var); }
inline Counted operator*(Counted a, Counted b) { count(C_MUL, a.var || b.var); return Counted::make(a.v * b.v, a.var || b.var); }
inline Counted operator/(Counted a, Counted b) { count(C_DIV, a.var || b.var); return Counted::make(a.v / b.v, a.var || b.var); }
inline Counted &operator+=(Counted &a, Counted b) { return a = a + b; }
inline Counted &operator-=(Counted &a, Counted b) { return a = a - b; }
inline Counted &operator*=(Counted &a, Counted b) { return a = a * b; }
inline Counted &operator/=(Counted &a, Counted b) { return a = a / b; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

// free operations (found by argument-dependent lookup)
inline Counted fabs(Counted a) { return Counted::make(std::fabs(a.v), a.var); }
inline Counted fmax(Counted a, Counted b) { return Counted::make(std::fmax(a.v, b.v), a.var || b.var); }
inline Counted fmin(Counted a, Counted b) { return Counted::make(std::fmin(a.v, b.v), a.var || b.var); }
inline Counted nextafter(Counted a, Counted b) { return Counted::make(std::nextafter(a.v, b.v), a.var || b.var); }
inline Counted fma(Counted a, Counted b, Counted c)
{
    const bool var = a.var || b.var || c.var;
    count(C_FMA, var);
    return Counted::make(std::fma(a.v, b.v, c.v), var);
}
inline Counted sqrt(Counted a) { count(C_SQRT, a.var); return Counted::make(std::sqrt(a.v), a.var); }
inline Counted cos(Counted a) { count(C_OTHER, a.var); return Counted::make(std::cos(a.v), a.var); }
inline Counted sin(Counted a) { count(C_OTHER, a.var); return Counted::make(std::sin(a.v), a.var); }
inline Counted tanh(Counted a) { count(C_OTHER, a.var); return Counted::make(std::tanh(a.v), a.var); }

// state, time and step size loaded into a member vary from then on
inline void mark_varying(Counted &c) { c.var = true; }

// ---- attribution hooks of the oracle
inline void trial_begin() { st().mark = st().now; }
inline void trial_end(int outcome)
{
    State &s = st();
    Tally &b = s.bucket[s.phase][outcome];
    for (int k = 0; k < N_COUNT; ++k)
        b.c[k] += s.now.c[k] - s.mark.c[k];
    ++s.trials[s.phase][outcome];
    s.mark = s.now;
}
inline void close_phase()
{
    State &s = st();
    Tally &all = s.bucket[s.phase][B_ALL];
    for (int k = 0; k < N_COUNT; ++k)
        all.c[k] += s.now.c[k];
    s.now.clear();
    s.mark.clear();
}
inline void set_phase(int p)
{
    close_phase();
    st().phase = p;
}
inline void reset()
{
    State &s = st();
    std::memset(&s.bucket, 0, sizeof(s.bucket));
    std::memset(&s.trials, 0, sizeof(s.trials));
    s.now.clear();
    s.mark.clear();
    s.phase = PH_TRANSIENT;
}

}  // namespace fc

// deterministic elementary functions of the right-hand side and controller
namespace cm {
inline fc::Counted exp(fc::Counted a) { fc::count(fc::C_EXP, a.var); return fc::Counted::make(cm::exp(a.v), a.var); }
inline fc::Counted log(fc::Counted a) { fc::count(fc::C_LOG, a.var); return fc::Counted::make(cm::log(a.v), a.var); }
inline fc::Counted pow(fc::Counted a, fc::Counted b)
{
    const bool var = a.var || b.var;
    fc::count(fc::C_POW, var);
    return fc::Counted::make(cm::pow(a.v, b.v), var);
}
inline fc::Counted pown(fc::Counted a, int n)
{
    fc::count(fc::C_POWN, a.var);
    return fc::Counted::make(cm::pown(a.v, n), a.var);
}
template <int N>
inline fc::Counted root(fc::Counted a)  // counted with pow: the controller's (rtol/err)^(1/(p+1))
{
    fc::count(fc::C_POW, a.var);
    return fc::Counted::make(cm::root<N>(a.v), a.var);
}
}  // namespace cm

#define CLODE_ORACLE_TRIAL_BEGIN() fc::trial_begin()
#define CLODE_ORACLE_TRIAL_END(outcome) fc::trial_end(outcome)
#define CLODE_ORACLE_PHASE(p) fc::set_phase(p)
