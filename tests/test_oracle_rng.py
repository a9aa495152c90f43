"""Pins the oracle's noise stream: xoroshiro128+ (Blackman & Vigna 2016,
rotations 55/14/36; clode/cpp/clODE_random.cl:93-113), the seeding rule of
clode/cpp/CLODE.cpp:447-464 (word k = seed + k in int arithmetic) and the
polar normal generator (clODE_random.cl:150-183)."""
import numpy as np
from scipy import stats

from oracle import OracleEnsemble, randn_stream, xoroshiro_stream

M64 = (1 << 64) - 1


def xoroshiro_py(s0, s1, n):
    """Independent restatement of the published algorithm in Python integers."""

    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    out = []
    for _ in range(n):
        out.append((s0 + s1) & M64)
        s1 ^= s0
        s0, s1 = rotl(s0, 55) ^ s1 ^ ((s1 << 14) & M64), rotl(s1, 36)
    return out


def test_xoroshiro_matches_independent_restatement():
    for s0, s1 in [(1, 2), (0x0123456789ABCDEF, 0xFEDCBA9876543210), (12345, 12345 + 65536)]:
        got = xoroshiro_stream(s0, s1, 1000)
        assert [int(v) for v in got] == xoroshiro_py(s0, s1, 1000)


def test_xoroshiro_known_prefix():
    # state (1, 2): first outputs 3, then (rotl(1,55) ^ 3 ^ (3<<14)) + rotl(3,36)
    s0, s1 = 1, 2
    first = xoroshiro_stream(s0, s1, 2)
    assert int(first[0]) == 3
    n0 = ((1 << 55) ^ 3 ^ (3 << 14)) & M64
    n1 = (3 << 36) & M64
    assert int(first[1]) == (n0 + n1) & M64


def test_seed_rule_word_k_is_seed_plus_k():
    o = OracleEnsemble("decay", 1, 1, stepper="seuler", kind="base")
    n = 5
    o.set_problem_data(np.ones((1, n)), np.ones((1, n)))
    for seed in (0, 7, -3):
        o.seed_rng(seed)
        words = o.get_rng()
        expect = [(seed + k) & M64 for k in range(2 * n)]  # sign-extended int
        assert [int(w) for w in words] == expect
    # member i gets (seed + i, seed + nPts + i)
    o.seed_rng(100)
    w = o.get_rng()
    assert int(w[3]) == 103 and int(w[n + 3]) == 100 + n + 3


def test_polar_normal_statistics():
    for prec in ("double", "single"):
        z = randn_stream(42, 4242, 200000, precision=prec)
        assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
        assert stats.kstest(z, "norm").pvalue > 1e-3


def test_polar_pairs_and_rejection_rate():
    # the polar method draws pairs; 2 normals consume 2 uniforms per accepted try,
    # acceptance probability pi/4
    s0, s1 = 9, 99
    z = randn_stream(s0, s1, 20000)
    # replay the uniform stream to count tries
    words = xoroshiro_stream(s0, s1, 40000)
    u = words.astype(np.float64) * 2.0 ** -64
    tries = accepted = 0
    k = 0
    while accepted < 10000:
        x1, x2 = 2 * u[k] - 1, 2 * u[k + 1] - 1
        k += 2
        tries += 1
        w = x1 * x1 + x2 * x2
        if w < 1:
            # first output of the pair
            y = x1 * np.sqrt(-2 * np.log(w) / w)
            assert abs(y - z[2 * accepted]) < 1e-12 * max(1, abs(y))
            accepted += 1
    assert abs(accepted / tries - np.pi / 4) < 0.02
