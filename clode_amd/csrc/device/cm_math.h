// Deterministic elementary functions for the ensemble integrator.
//
// The integrator's results depend on exp/log/pow at three places: the user's
// right-hand side (e.g. Boltzmann gating functions), the adaptive step-size
// controller (pow of the error ratio) and the polar Gaussian generator (log).
// Vendor math libraries differ by an ulp or two between the host libm and the
// GPU's device library, which would make FP64 runs differ in the last bits from
// a CPU restatement of the same algorithm.  These routines are written with
// explicit fused multiply-adds, exactly rounded IEEE operations (+,-,*,/,
// sqrt, rint, ldexp, frexp) and fixed constants only, so they return the same
// bits wherever they are compiled (hipRTC for gfx950, clang++ for x86-64 with
// FMA).  Accuracy against glibc is checked in the test-suite (<= 2 ulp for
// exp, log and pow; the OpenCL C 1.2 bounds are 3, 3 and 16 ulp).
//
// Methods:
//  exp: Cody-Waite argument reduction x = k ln2 + r, |r| <= ln2/2, followed by
//       e^r = 1 + r + r^2 Q(r) with a fitted polynomial Q (degree 9 for double,
//       4 for float) in Horner form, scaled exactly by 2^k with ldexp.
//  log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f);
//       log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))) where R is the series
//       2z/3 + 2z^2/5 + ... of 2 atanh(s)/s - 2 (11 terms double, 4 float).
//  pow: IEEE-754 special cases, then exp(p) (1 + q) where p + q = y log|x| is
//       formed in double-double (log from the atanh series with exact two-sums
//       and fma-based products), with the sign of x for odd integer y; the
//       float version rounds the double result once.
//  pown: binary exponentiation (exact operation sequence).
//  root<N>: x^(1/N), N = 3, 5, from a polynomial guess and Newton steps (below);
//           N = 2 is the square root.
//
// CM_FN is defined by the including translation unit (host or device).
#ifndef CLODE_CM_MATH_H
#define CLODE_CM_MATH_H

#ifndef CM_FN
#define CM_FN inline
#endif

namespace cm {

// ---------------------------------------------------------------- double --
CM_FN double exp(double x)
{
    // Branch-free: the reduction runs on x clamped to [-746, 710] (NaN clamps
    // to -746) and NaN is selected at the end.  Overflow and underflow need no
    // select: above ln(DBL_MAX) the reduced r is positive, so 2^k p with
    // k = 1024 and p > 1 overflows to inf, and below -1075 ln 2 it is
    // negative, so 2^-1075 p with p < 1 rounds to 0.
    const double log2e = 0x1.71547652b82fep+0;
    const double ln2hi = 0x1.62e42fefa39efp-1;
    const double ln2lo = 0x1.abc9e3b39803fp-56;
    double xc = fmin(fmax(x, -746.0), 710.0);
    double k = rint(xc * log2e);
    double r = fma(-k, ln2hi, xc);
    r = fma(-k, ln2lo, r);
    // exp(r) = 1 + r + r^2 Q(r) on |r| <= ln2/2: Q of degree 9 interpolating
    // (e^r - 1 - r)/r^2 at the Chebyshev nodes (|error| 1.1e-16, relative
    // error of exp 1.7e-17)
    double p = 0x1.af38a57235af4p-26;
    p = fma(p, r, 0x1.28917e8165204p-22);
    p = fma(p, r, 0x1.71de0db19876ap-19);
    p = fma(p, r, 0x1.a019b912dc201p-16);
    p = fma(p, r, 0x1.a01a01a7c2483p-13);
    p = fma(p, r, 0x1.6c16c1788a61fp-10);
    p = fma(p, r, 0x1.11111111109b9p-7);
    p = fma(p, r, 0x1.5555555553d67p-5);
    p = fma(p, r, 0x1.5555555555556p-3);
    p = fma(p, r, 0x1.0000000000001p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    double y = ldexp(p, (int)k);
    return (x == x) ? y : x;
}

CM_FN double log(double x)
{
    // Branch-free: the reduction runs on x, or on 1 where x is not a positive
    // finite number, and the IEEE special cases are selected at the end.
    double xs = (x > 0.0 && x < __builtin_huge_val()) ? x : 1.0;
    int e;
    double m = frexp(xs, &e);  // xs = m 2^e, m in [0.5, 1)
    bool lowm = m < 0x1.6a09e667f3bcdp-1;
    m = lowm ? m * 2.0 : m;
    e = lowm ? e - 1 : e;
    const double ln2hi = 0x1.62e42fee00000p-1;  // 21 trailing zero bits: e*ln2hi exact
    const double ln2lo = 0x1.a39ef35793c76p-33;
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0x1.642c8590b2164p-4;
    R = fma(R, z, 0x1.8618618618618p-4);
    R = fma(R, z, 0x1.af286bca1af28p-4);
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
    R = fma(R, z, 0x1.1111111111111p-3);
    R = fma(R, z, 0x1.3b13b13b13b14p-3);
    R = fma(R, z, 0x1.745d1745d1746p-3);
    R = fma(R, z, 0x1.c71c71c71c71cp-3);
    R = fma(R, z, 0x1.2492492492492p-2);
    R = fma(R, z, 0x1.999999999999ap-2);
    R = fma(R, z, 0x1.5555555555555p-1);
    R = R * z;
    double hfsq = (0.5 * f) * f;
    double de = (double)e;
    double lo = fma(s, hfsq + R, de * ln2lo);
    double y = de * ln2hi - ((hfsq - lo) - f);
    y = (x == __builtin_huge_val()) ? x : y;
    y = (x == 0.0) ? -__builtin_huge_val() : y;
    return (x < 0.0 || !(x == x)) ? __builtin_nan("") : y;
}

// log(x) as an unevaluated sum hi + lo (|lo| <= ulp(hi)), x finite and > 0:
// log x = e ln2 + 2 atanh(s), s = (m-1)/(m+1), with s, m+1 and the leading
// terms carried in double-double (Dekker products / Knuth two-sums via fma).
CM_FN double log_dd(double x, double *lo)
{
    int e;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e = e - 1;
    }
    double num = m - 1.0;             // exact
    double den = m + 1.0;             // two-sum: den + den_lo = m + 1 exactly
    double dbb = den - m;
    double den_lo = (m - (den - dbb)) + (1.0 - dbb);
    double s = num / den;
    double r = fma(-s, den, num);     // num - s*den exactly
    r = fma(-s, den_lo, r);
    double s_lo = r / den;
    double z = s * s;
    double T = 0x1.642c8590b2164p-4;
    T = fma(T, z, 0x1.8618618618618p-4);
    T = fma(T, z, 0x1.af286bca1af28p-4);
    T = fma(T, z, 0x1.e1e1e1e1e1e1ep-4);
    T = fma(T, z, 0x1.1111111111111p-3);
    T = fma(T, z, 0x1.3b13b13b13b14p-3);
    T = fma(T, z, 0x1.745d1745d1746p-3);
    T = fma(T, z, 0x1.c71c71c71c71cp-3);
    T = fma(T, z, 0x1.2492492492492p-2);
    T = fma(T, z, 0x1.999999999999ap-2);
    T = fma(T, z, 0x1.5555555555555p-1);
    T = T * z * s;                    // 2 atanh(s) = 2s + T
    const double ln2hi = 0x1.62e42fee00000p-1;
    const double ln2lo = 0x1.a39ef35793c76p-33;
    double de = (double)e;
    double a = de * ln2hi;            // exact
    double b = 2.0 * s;               // exact
    double hi = a + b;                // two-sum
    double bb = hi - a;
    double err = (a - (hi - bb)) + (b - bb);
    double l = err + 2.0 * s_lo + T + de * ln2lo;
    double h2 = hi + l;               // renormalise
    *lo = l - (h2 - hi);
    return h2;
}

CM_FN double pow(double x, double y)
{
    // Branch-free: the double-double path runs on a sanitised |x| and the
    // IEEE-754 special cases are selected afterwards, highest priority last.
    double ax = fabs(x);
    double ay = fabs(y);
    bool yint = (y == rint(y));
    bool yodd = yint && ay < 0x1.0p53 && (0.5 * y != rint(0.5 * y));
    bool neg = x < 0.0 && yodd;
    double axs = (ax > 0.0 && ax < __builtin_huge_val()) ? ax : 1.0;
    double ys = (ay < __builtin_huge_val()) ? y : 0.0;  // also NaN -> 0
    // y log|x| in double-double, then exp(p + q) = exp(p) (1 + q)
    double llo;
    double lhi = cm::log_dd(axs, &llo);
    double p = ys * lhi;
    double q = fma(ys, lhi, -p);
    q = fma(ys, llo, q);
    double r = cm::exp(p);
    r = (r != 0.0 && r != __builtin_huge_val()) ? fma(r, q, r) : r;
    r = neg ? -r : r;
    const double inf = __builtin_huge_val();
    double rinf = (y > 0.0) ? inf : 0.0;                 // |x| = inf
    r = (ax == inf) ? (neg ? -rinf : rinf) : r;
    double rzero = (y > 0.0) ? (yodd ? x : 0.0) : (yodd ? __builtin_copysign(inf, x) : inf);
    r = (ax == 0.0) ? rzero : r;                         // x = +-0
    r = (x < 0.0 && !yint) ? __builtin_nan("") : r;
    double ryinf = (ax == 1.0) ? 1.0 : (((ax > 1.0) == (y > 0.0)) ? inf : 0.0);
    r = (ay == inf) ? ryinf : r;                         // y = +-inf
    r = (!(x == x) || !(y == y)) ? x + y : r;           // NaN
    return (y == 0.0 || x == 1.0) ? 1.0 : r;
}

CM_FN double pown(double x, int n)
{
    unsigned int m = n < 0 ? (unsigned int)(-(long long)n) : (unsigned int)n;
    double r = 1.0;
    double b = x;
    while (m) {
        if (m & 1u)
            r = r * b;
        b = b * b;
        m >>= 1;
    }
    return n < 0 ? 1.0 / r : r;
}

// x^(1/N) for N = 3 or 5: the adaptive controller's factor (rtol/err)^(1/(p+1))
// for the error orders p = 2 (bs23) and 4 (dopri5) (N = 2, p = 1 for he12, is
// the IEEE square root, which is what pow(x, 0.5) rounds to), with the special cases of
// pow(x, 1/N): NaN for NaN or x < 0, +0 for +-0, inf for inf.  Branch-free and
// division-free:
//   x = M 2^(N q) with M = m 2^r, m in [1/2, 1), 0 <= r < N;
//   w ~ M^(-1/N) as P(m) S(r): P the minimax quintic of m^(-1/N) on [1/2, 1)
//   (relative error 2.7e-6 for N = 5, 5.7e-6 for N = 3) and S the polynomial
//   of degree N-1 through the N points (r, 2^(-r/N)); one Newton step
//   w <- w + w (1 - M w^N) / N (relative error 3 eps^2 or 2 eps^2);
//   y = M w^(N-1) and one Newton step on y^N = M, scaled by 2^q.
// Within 1 ulp of the exact root (checked in the test-suite against
// long-double references); it differs from pow(x, RN(1/N)) by the exponent's
// own rounding, |log x| 2^-55 relative.
template <int N>
CM_FN double root(double x)
{
    static_assert(N == 2 || N == 3 || N == 5, "root<N>: N = 2, 3 or 5");
    if (N == 2)  // pow(x, 0.5) is the correctly rounded square root (+0 for -0)
        return x == 0.0 ? 0.0 : sqrt(x);
    const double inf = __builtin_huge_val();
    const bool finitePos = x > 0.0 && x < inf;
    double xs = finitePos ? x : 1.0;
    int e;
    double m = frexp(xs, &e);  // xs = m 2^e, m in [0.5, 1), e in [-1073, 1024]
    // floor(e / N) and e mod N without a sign test: 1080 is a multiple of 3
    // and of 5 and keeps e + 1080 positive
    const unsigned eb = (unsigned)(e + 1080), qb = eb / N;
    const int q = (int)qb - 1080 / N;
    const int r = (int)(eb - N * qb);  // 0 .. N-1
    double M = ldexp(m, r);
    const double rd = (double)r;
    double w, sc;
    if (N == 3) {
        w = fma(fma(fma(fma(fma(-0x1.5f50df22b3948p-1, m, 0x1.92b6d1f866c72p+1), m, -0x1.81aceb53d20cdp+2), m,
                        0x1.912b993c1c0b2p+2), m, -0x1.f91b483312656p+1), m, 0x1.1f3b21f958b0dp+1);
        sc = fma(fma(0x1.5ca5af129611ep-6, rd, -0x1.d2150c92f7332p-3), rd, 1.0);  // 2^(-r/3)
    } else {
        w = fma(fma(fma(fma(fma(-0x1.3fec119f23373p-2, m, 0x1.706cb0760b0d0p+0), m, -0x1.63bf8a58843cbp+1), m,
                        0x1.784027ecd4f80p+1), m, -0x1.ef99a9737ac22p+0), m, 0x1.a62695914b34dp+0);
        sc = fma(fma(fma(fma(0x1.88970ec6f3ac2p-17, rd, -0x1.c4b4e3b1a9127p-12), rd, 0x1.3a4e88f4a3e7ap-7), rd,
                     -0x1.1be579c2b43dbp-3), rd, 1.0);  // 2^(-r/5)
    }
    w *= sc;
    // one Newton step for w = M^(-1/N): relative error 6e-6 -> 7e-11
    const double rN = 1.0 / N;  // folded: RN(1/N)
    {
        double w2 = w * w;
        double wN = (N == 3) ? w2 * w : (w2 * w2) * w;
        double t = fma(-M, wN, 1.0);
        w = fma(w * t, rN, w);
    }
    double w2 = w * w;
    double wN1 = (N == 3) ? w2 : w2 * w2;  // w^(N-1)
    double y = M * wN1;
    double y2 = y * y;
    double yN = (N == 3) ? y2 * y : (y2 * y2) * y;
    y = fma(-(yN - M) * wN1, rN, y);
    y = ldexp(y, q);
    // +-0 -> +0 and inf -> inf (x + 0 is +0 for -0); NaN and x < 0 -> NaN
    double special = (x >= 0.0) ? x + 0.0 : __builtin_nan("");
    return finitePos ? y : special;
}

// float: the double root, rounded once
template <int N>
CM_FN float root(float x)
{
    return (float)root<N>((double)x);
}

// The adaptive controller's minimum step, 16 |fabs(nextafter(t, 1.1 t1)) - t|
// (1.1 t1 given as `toward`), exactly.  For a normal t > 0 below the largest
// finite value, stepping toward a larger value, nextafter adds the spacing of
// the binade, 2^(ilogb(t) - 52), so the result is 2^(ilogb(t) - 48): an
// exponent read and a scale instead of nextafter's integer sequence.  Other
// arguments (t <= 0, subnormal, the largest finite value, a step downward,
// NaN) take the formula itself.
CM_FN double min_step(double t, double toward)
{
    if (t >= 0x1p-1022 && t < 0x1.fffffffffffffp+1023 && t < toward) {
        int e;
        (void)frexp(t, &e);  // t = m 2^e, m in [1/2, 1): ilogb(t) = e - 1
        return ldexp(1.0, e - 49);
    }
    return 16.0 * fabs(fabs(nextafter(t, toward)) - t);
}

// float: spacing 2^(ilogb(t) - 23), so 16 spacings are 2^(ilogb(t) - 19)
CM_FN float min_step(float t, float toward)
{
    if (t >= 0x1p-126f && t < 0x1.fffffep+127f && t < toward) {
        int e;
        (void)frexpf(t, &e);
        return ldexpf(1.0f, e - 20);
    }
    return 16.0f * fabsf(fabsf(nextafterf(t, toward)) - t);
}

// ----------------------------------------------------------------- float --
CM_FN float exp(float x)
{
    // branch-free, as the double version (overflow above ln(FLT_MAX) with
    // k = 128, underflow to 0 below -150 ln 2)
    const float log2e = 0x1.715476p+0f;
    const float ln2hi = 0x1.62e430p-1f;
    const float ln2lo = -0x1.05c610p-29f;
    float xc = fminf(fmaxf(x, -104.0f), 89.0f);
    float k = rintf(xc * log2e);
    float r = fmaf(-k, ln2hi, xc);
    r = fmaf(-k, ln2lo, r);
    // exp(r) = 1 + r + r^2 Q(r): Q of degree 4, minimax in relative error
    // (relative error of exp 9e-9)
    float p = 0x1.6c69f2p-10f;
    p = fmaf(p, r, 0x1.120b1cp-7f);
    p = fmaf(p, r, 0x1.5555bap-5f);
    p = fmaf(p, r, 0x1.5554dep-3f);
    p = fmaf(p, r, 0x1.000000p-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    float y = ldexpf(p, (int)k);
    return (x == x) ? y : x;
}

CM_FN float log(float x)
{
    // branch-free, as the double version
    float xs = (x > 0.0f && x < __builtin_huge_valf()) ? x : 1.0f;
    int e;
    float m = frexpf(xs, &e);
    bool lowm = m < 0x1.6a09e6p-1f;
    m = lowm ? m * 2.0f : m;
    e = lowm ? e - 1 : e;
    const float ln2hi = 0x1.62e400p-1f;  // 15 trailing zero bits: e*ln2hi exact
    const float ln2lo = 0x1.7f7d1cp-20f;
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 0x1.c71c72p-3f;
    R = fmaf(R, z, 0x1.24924ap-2f);
    R = fmaf(R, z, 0x1.99999ap-2f);
    R = fmaf(R, z, 0x1.555556p-1f);
    R = R * z;
    float hfsq = (0.5f * f) * f;
    float de = (float)e;
    float lo = fmaf(s, hfsq + R, de * ln2lo);
    float y = de * ln2hi - ((hfsq - lo) - f);
    y = (x == __builtin_huge_valf()) ? x : y;
    y = (x == 0.0f) ? -__builtin_huge_valf() : y;
    return (x < 0.0f || !(x == x)) ? __builtin_nanf("") : y;
}

CM_FN float pow(float x, float y)
{
    if (y == 0.0f || x == 1.0f)
        return 1.0f;
    if (!(x == x) || !(y == y))
        return x + y;
    float ax = fabsf(x);
    float ay = fabsf(y);
    bool yint = (y == rintf(y));
    bool yodd = yint && ay < 0x1.0p24f && (0.5f * y != rintf(0.5f * y));
    if (ay == __builtin_huge_valf()) {
        if (ax == 1.0f)
            return 1.0f;
        return ((ax > 1.0f) == (y > 0.0f)) ? __builtin_huge_valf() : 0.0f;
    }
    if (x < 0.0f && !yint)
        return __builtin_nanf("");
    if (ax == 0.0f) {
        if (y > 0.0f)
            return yodd ? x : 0.0f;
        return yodd ? __builtin_copysignf(__builtin_huge_valf(), x) : __builtin_huge_valf();
    }
    if (ax == __builtin_huge_valf()) {
        float r = (y > 0.0f) ? __builtin_huge_valf() : 0.0f;
        return (x < 0.0f && yodd) ? -r : r;
    }
    // evaluated in double: one final rounding to float
    float r = (float)cm::pow((double)ax, (double)y);
    return (x < 0.0f && yodd) ? -r : r;
}

CM_FN float pown(float x, int n)
{
    unsigned int m = n < 0 ? (unsigned int)(-(long long)n) : (unsigned int)n;
    float r = 1.0f;
    float b = x;
    while (m) {
        if (m & 1u)
            r = r * b;
        b = b * b;
        m >>= 1;
    }
    return n < 0 ? 1.0f / r : r;
}

}  // namespace cm

#endif  // CLODE_CM_MATH_H
