// OpenCL C built-ins that a user's getRHS may call and that HIP/C++ lack.
//
// The right-hand-side source is written in the OpenCL C dialect accepted by
// clODE (clode/function_converter.py:55-125 lists the built-ins its Python
// front-end emits; clode/cpp/clODE_utilities.cl:17-19 adds MIN, MAX and
// heaviside).  This header is compiled both into hipRTC device programs and
// into the host-side CPU restatement, so CM_FN is defined by the includer.
// exp, log, pow and pown resolve to the deterministic routines of cm_math.h
// unless CLODE_PLATFORM_MATH is defined (the fp32-compensated mode uses the
// platform's fast single-precision library instead).
#ifndef CLODE_BUILTINS_H
#define CLODE_BUILTINS_H

#include "cm_math.h"

namespace cm {
// Mixed-type calls follow C's usual arithmetic conversions (promote to double).
template <class A, class B>
CM_FN double pow(A a, B b) { return cm::pow((double)a, (double)b); }
template <class A>
CM_FN double exp(A a) { return cm::exp((double)a); }
template <class A>
CM_FN double log(A a) { return cm::log((double)a); }
template <class A>
CM_FN double pown(A x, int n) { return cm::pown((double)x, n); }
}  // namespace cm

#ifndef CLODE_PLATFORM_MATH
#define exp(x) cm::exp(x)
#define log(x) cm::log(x)
#define pow(x, y) cm::pow(x, y)
#endif
#define pown(x, n) cm::pown(x, n)

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define heaviside(x) ((x) >= ZERO ? ONE : ZERO)

#define CLODE_PI_D 3.141592653589793238462643383279502884
#define CLODE_PI_F 3.14159265358979323846f

CM_FN double rootn(double x, int n)
{
    if (n == 0)
        return __builtin_nan("");
    if (x < 0.0) {
        if ((n & 1) == 0)
            return __builtin_nan("");
        return -pow(-x, 1.0 / (double)n);
    }
    return pow(x, 1.0 / (double)n);
}
CM_FN float rootn(float x, int n)
{
    if (n == 0)
        return __builtin_nanf("");
    if (x < 0.0f) {
        if ((n & 1) == 0)
            return __builtin_nanf("");
        return -pow(-x, 1.0f / (float)n);
    }
    return pow(x, 1.0f / (float)n);
}
CM_FN double powr(double x, double y) { return x < 0.0 ? __builtin_nan("") : pow(x, y); }
CM_FN float powr(float x, float y) { return x < 0.0f ? __builtin_nanf("") : pow(x, y); }

CM_FN double tanpi(double x) { return tan(CLODE_PI_D * x); }
CM_FN float tanpi(float x) { return tanf(CLODE_PI_F * x); }
CM_FN double asinpi(double x) { return asin(x) / CLODE_PI_D; }
CM_FN float asinpi(float x) { return asinf(x) / CLODE_PI_F; }
CM_FN double acospi(double x) { return acos(x) / CLODE_PI_D; }
CM_FN float acospi(float x) { return acosf(x) / CLODE_PI_F; }
CM_FN double atanpi(double x) { return atan(x) / CLODE_PI_D; }
CM_FN float atanpi(float x) { return atanf(x) / CLODE_PI_F; }
CM_FN double atan2pi(double y, double x) { return atan2(y, x) / CLODE_PI_D; }
CM_FN float atan2pi(float y, float x) { return atan2f(y, x) / CLODE_PI_F; }

CM_FN double clamp(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
CM_FN float clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
CM_FN double mad(double a, double b, double c) { return a * b + c; }
CM_FN float mad(float a, float b, float c) { return a * b + c; }
CM_FN double sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0 * x); }
CM_FN float sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f * x); }
CM_FN double step(double edge, double x) { return x < edge ? 0.0 : 1.0; }
CM_FN float step(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
CM_FN double degrees(double r) { return r * (180.0 / CLODE_PI_D); }
CM_FN float degrees(float r) { return r * (180.0f / CLODE_PI_F); }
CM_FN double radians(double d) { return d * (CLODE_PI_D / 180.0); }
CM_FN float radians(float d) { return d * (CLODE_PI_F / 180.0f); }

#endif  // CLODE_BUILTINS_H
