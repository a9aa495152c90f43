// Per-member pseudo-random stream: xoroshiro128+ (Blackman & Vigna, 2016;
// rotation constants a=55, b=14, c=36) with a uniform [0,1) draw and a
// Marsaglia polar-method normal draw that caches its second value.
//
// Follows clode/cpp/clODE_random.cl:93-183: the 64-bit output s0+s1 is
// converted to realtype and scaled by 2^-64 (so in single precision the draw
// can round up to 1.0f, exactly as the reference does), the polar loop rejects
// w >= 1, and the cache flag is reset at every kernel launch by the kernel
// (clode/cpp/transient.cl:42).  The 64-bit stream is therefore identical bit
// for bit to the reference's for the same seed words.
#ifndef CLODE_RNG_H
#define CLODE_RNG_H

#define CLODE_RNGNORM RCONST(5.421010862427522e-20)
#define CLODE_POLAR_MAX_TRIES 4096

struct RngData {
    unsigned long long s0;
    unsigned long long s1;
    realtype last;
    bool useLast;
};

CLODE_DEV unsigned long long rng_rotl(unsigned long long x, int k)
{
    return (x << k) | (x >> (64 - k));
}

CLODE_DEV unsigned long long rng_next(RngData &rd)
{
    const unsigned long long s0 = rd.s0;
    unsigned long long s1 = rd.s1;
    const unsigned long long result = s0 + s1;
    s1 ^= s0;
    rd.s0 = rng_rotl(s0, 55) ^ s1 ^ (s1 << 14);
    rd.s1 = rng_rotl(s1, 36);
    return result;
}

CLODE_DEV realtype rng_uniform(RngData &rd)
{
    unsigned long long r = rng_next(rd);
    return (realtype)r * CLODE_RNGNORM;
}

CLODE_DEV realtype rng_normal(RngData &rd)
{
    realtype y1;
    if (rd.useLast) {
        y1 = rd.last;
        rd.useLast = false;
    } else {
        realtype x1, x2, w;
        int guard = 0;
        // bounded: only the degenerate all-zero state never leaves the loop
        do {
            x1 = RCONST(2.0) * rng_uniform(rd) - ONE;
            x2 = RCONST(2.0) * rng_uniform(rd) - ONE;
            w = x1 * x1 + x2 * x2;
        } while (w >= ONE && ++guard < CLODE_POLAR_MAX_TRIES);
        w = sqrt((-RCONST(2.0) * log(w)) / w);
        y1 = x1 * w;
        rd.last = x2 * w;
        rd.useLast = true;
    }
    return y1;
}

#endif  // CLODE_RNG_H
