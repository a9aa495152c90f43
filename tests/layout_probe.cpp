// Host compile of the device observer header (clode_amd/csrc/device/
// clode_observers.h) that counts the persistent fields an observer spills to
// HBM and compares them with the host-side layout (clode_layout.h), which
// sizes the buffers.  Built and run by tests/test_native_cpu.py.
#include <cmath>
#include <cstddef>
#include <cstdio>

#define CLODE_DEV inline
#define CM_FN inline
#define CLODE_PREC 2
typedef double realtype;
typedef double timetype;
#define RCONST(x) (x)
#define BIG_REAL 1e300
#define ZERO 0.0
#define ONE 1.0
#define N_PAR1 1
#define N_AUX1 (N_AUX > 0 ? N_AUX : 1)
using std::fmax;
using std::fmin;
using std::sqrt;

#include "../clode_amd/csrc/device/clode_observers.h"

struct Counter {
    int r = 0, u = 0;
    template <class T>
    void real(T &) { r++; }
    void uint(unsigned &) { u++; }
};

int main()
{
    Observer ob;
    Counter c;
    ob.fields(c);
    std::printf("real %d (layout %d), uint %d (layout %d)\n", c.r, Observer::NSCALAR, c.u,
                clode_obs_nuint(CLODE_OBSERVER));
    return !(c.r == Observer::NSCALAR && c.u == clode_obs_nuint(CLODE_OBSERVER));
}
