// Prelude of every hipRTC program: working precision, the OpenCL C dialect
// names used by right-hand-side sources, and the deterministic math routines.
//
// Precision modes (CLODE_PREC):
//   1  single  : realtype = float,  time in float  (clode/cpp/realtype.cl:9-21)
//   2  double  : realtype = double, time in double (clode/cpp/realtype.cl:23-45)
//   3  fp32c   : realtype = float for the state, the right-hand side and
//                stage algebra; time, step size, error norm, observer
//                statistics and continuation state in double; state updates
//                Kahan-compensated.  Uses the platform's single-precision
//                math library (exp/log/pow) and division for speed, with
//                float denormals flushed to zero, divisions allowed to
//                multiply by a reciprocal and math functions allowed to use
//                the hardware approximations.
#ifndef CLODE_PRELUDE_H
#define CLODE_PRELUDE_H

#define CLODE_DEV __device__ __attribute__((always_inline)) inline
#define CM_FN __host__ __device__ __attribute__((always_inline)) inline

#ifndef CLODE_PREC
#error "CLODE_PREC must be defined (1 single, 2 double, 3 fp32c)"
#endif

#if CLODE_PREC == 3
#define CLODE_PLATFORM_MATH 1
#endif

#if CLODE_PREC == 2
typedef double realtype;
typedef double timetype;
#define RCONST(x) (x)
#define BIG_REAL 1.7976931348623157e308
#define SMALL_REAL 2.2250738585072014e-308
#define UNIT_ROUNDOFF 2.220446049250313e-16
#define ZERO 0.0
#define ONE 1.0
#else
typedef float realtype;
#if CLODE_PREC == 1
typedef float timetype;
#else
typedef double timetype;
#endif
#define RCONST(x) (x##f)
#define BIG_REAL 3.40282347e+38f
#define SMALL_REAL 1.17549435e-38f
#define UNIT_ROUNDOFF 1.1920929e-07f
#define ZERO 0.0f
#define ONE 1.0f
#endif

#ifndef FLT_MAX
#define FLT_MAX 3.40282347e+38f
#define FLT_MIN 1.17549435e-38f
#define FLT_EPSILON 1.1920929e-07f
#endif
#ifndef DBL_MAX
#define DBL_MAX 1.7976931348623157e308
#define DBL_MIN 2.2250738585072014e-308
#define DBL_EPSILON 2.220446049250313e-16
#endif

// OpenCL C address-space and access qualifiers that may appear in RHS sources
#ifndef __private
#define __private
#endif
#ifndef __global
#define __global
#endif
#ifndef __constant
#define __constant const
#endif

#include "cm_math.h"
#include "clode_builtins.h"

#endif  // CLODE_PRELUDE_H
