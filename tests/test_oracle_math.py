"""Pins the deterministic elementary functions (clode_amd/csrc/device/cm_math.h)
shared by the GPU programs and the CPU oracle against glibc/numpy.

exp, log and pow must stay within 2 ulp of the correctly rounded value (the
OpenCL C 1.2 bounds are 3, 3 and 16 ulp); pown is exact repeated
multiplication.  IEEE special cases follow C99 Annex F.
"""
import numpy as np
import pytest

from oracle import cm_eval


def ulp_err64(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def ulp_err32(got, x):
    ref32 = np.float32(x)
    return np.abs(got - ref32.astype(np.float64)) / np.spacing(np.abs(ref32)).astype(np.float64)


def test_exp_double_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-700, 700, 200000), rng.uniform(-1, 1, 100000), rng.uniform(-1e-6, 1e-6, 1000)])
    got = cm_eval("exp", x)
    assert ulp_err64(got, np.exp(x)).max() <= 2.0


def test_exp_double_against_extended_precision_and_thresholds():
    """exp against x87 extended precision (64-bit significand, about 2^-11 ulp
    of double): within 1 ulp everywhere and correctly rounded for most
    arguments.  Around the overflow threshold ln(DBL_MAX) and the underflow
    threshold -1075 ln 2, which the reduction handles without selects, every
    double in dense sweeps gives numpy's result bit for bit (inf, DBL_MAX,
    subnormals, 0), so a dropped or misplaced clamp would show there."""
    assert np.finfo(np.longdouble).nmant >= 63
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-708, 709, 300000), rng.uniform(-1, 1, 100000), rng.uniform(-40, 40, 100000)])
    got = cm_eval("exp", x)
    ref = np.exp(x.astype(np.longdouble))
    err = np.abs((got.astype(np.longdouble) - ref) / np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble))
    assert float(err.max()) <= 1.0
    assert np.mean(err <= 0.5) > 0.9
    for c in (709.782712893384, -745.1332191019412, -708.3964185322641, -744.44007192138122):
        sweep = c + np.arange(-5000, 5000) * np.spacing(c)
        with np.errstate(over="ignore"):
            want = np.exp(sweep)
        assert np.array_equal(cm_eval("exp", sweep), want), c
    assert cm_eval("exp", np.array([-1075 * np.log(2)]))[0] in (0.0, 5e-324)
    # float: the same thresholds ln(FLT_MAX) and -150 ln 2; overflow and
    # flush to zero exactly where numpy's float32 exp has them, 2 ulp elsewhere
    for c in (88.72283935546875, -103.97208404541015625, -87.33654):
        c32 = np.float32(c)
        sweep = (c32 + np.arange(-3000, 3000, dtype=np.float32) * np.spacing(c32)).astype(np.float64)
        got = cm_eval("exp", sweep, precision="single")
        with np.errstate(over="ignore"):
            want = np.exp(sweep.astype(np.float32)).astype(np.float64)
        assert np.array_equal(np.isinf(got), np.isinf(want)) and np.array_equal(got == 0, want == 0), c
        fin = np.isfinite(want) & (want > 0)
        assert ulp_err32(got[fin], want[fin]).max() <= 2.0, c


def test_log_double_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([10 ** rng.uniform(-300, 300, 200000), rng.uniform(0.5, 2.0, 100000), 1 + rng.uniform(-1e-8, 1e-8, 1000)])
    got = cm_eval("log", x)
    ref = np.log(x)
    ok = np.abs(ref) > 0
    assert ulp_err64(got[ok], ref[ok]).max() <= 2.0


def test_exp_log_float_accuracy():
    rng = np.random.default_rng(3)
    x = rng.uniform(-87, 88, 200000).astype(np.float32).astype(np.float64)
    assert ulp_err32(cm_eval("exp", x, precision="single"), np.exp(x)).max() <= 2.0
    y = (10 ** rng.uniform(-37, 38, 200000)).astype(np.float32).astype(np.float64)
    ref = np.log(y)
    ok = np.abs(ref) > 1e-30
    assert ulp_err32(cm_eval("log", y, precision="single")[ok], ref[ok]).max() <= 2.0


def test_pow_accuracy():
    rng = np.random.default_rng(4)
    x = 10 ** rng.uniform(-5, 5, 100000)
    y = rng.uniform(-10, 10, 100000)
    ref = np.power(x, y)
    ok = np.isfinite(ref) & (ref > 1e-300) & (ref < 1e300)
    assert ulp_err64(cm_eval("pow", x, y)[ok], ref[ok]).max() <= 2
    # the controller's use: (rtol/err)^(1/5), (rtol/err)^(1/3)
    q = 10 ** rng.uniform(-8, 8, 10000)
    for e in (0.2, 1.0 / 3.0):
        assert ulp_err64(cm_eval("pow", q, np.full_like(q, e)), np.power(q, e)).max() <= 4


def test_special_cases():
    inf, nan = np.inf, np.nan
    x = np.array([0.0, -0.0, inf, -inf, nan, 710.0, -746.0])
    e = cm_eval("exp", x)
    assert e[0] == 1 and e[1] == 1 and e[2] == inf and e[3] == 0 and np.isnan(e[4]) and e[5] == inf and e[6] == 0
    lg = cm_eval("log", np.array([0.0, -1.0, inf, 1.0, nan]))
    assert lg[0] == -inf and np.isnan(lg[1]) and lg[2] == inf and lg[3] == 0 and np.isnan(lg[4])
    px = np.array([2.0, -2.0, -2.0, 0.0, 0.0, inf, 1.0, nan, -8.0])
    py = np.array([0.0, 3.0, 0.5, 2.0, -1.0, 0.5, nan, 0.0, 1.0 / 3.0])
    p = cm_eval("pow", px, py)
    assert p[0] == 1 and p[1] == -8 and np.isnan(p[2]) and p[3] == 0 and p[4] == inf and p[5] == inf
    assert p[6] == 1 and p[7] == 1 and np.isnan(p[8])


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 7, -1, -2, -3])
def test_pown_exact(n):
    x = np.array([1.5, -2.25, 0.3, 7.0])
    got = cm_eval("pown", x, np.full_like(x, n))
    ref = np.array([float(np.float64(v) ** n) for v in x])
    assert np.allclose(got, ref, rtol=4e-16, atol=0)


def root_reference(x, n):
    """x^(1/n) in x87 extended precision (64-bit significand), rounded once to
    double: within about 0.5 ulp of the exact root for every double x > 0."""
    xl = x.astype(np.longdouble)
    return np.exp(np.log(xl) / np.longdouble(n)).astype(np.float64)


@pytest.mark.parametrize("n", [3, 5])
def test_fixed_root_accuracy(n):
    """cm::root<N> (the controller's (rtol/err)^(1/(p+1))) against an
    extended-precision reference: within 1 ulp over the whole double range,
    subnormals included, and against numpy's cube root for N = 3; the IEEE
    special cases of pow(x, 1/N)."""
    assert np.finfo(np.longdouble).nmant >= 63
    rng = np.random.default_rng(40 + n)
    x = np.concatenate([10 ** rng.uniform(-307, 308, 200000), rng.uniform(0.5, 2.0, 50000),
                        2.0 ** rng.integers(-1074, 1023, 5000).astype(np.float64),
                        rng.uniform(1, 1000, 2000) * 5e-324])
    got = cm_eval(f"root{n}", x)
    err = ulp_err64(got, root_reference(x, n))
    assert err.max() <= 1.0, (n, float(err.max()), x[np.argmax(err)])
    assert np.mean(err <= 0.5) > 0.85  # correctly rounded for most x
    # pow(x, RN(1/N)) is further from the root: its exponent is rounded
    assert ulp_err64(np.power(x, 1.0 / n), root_reference(x, n)).max() > 10
    if n == 3:
        assert ulp_err64(got, np.cbrt(x)).max() <= 2.0
    # exact powers (k^N < 2^53): (k^N)^(1/N) = k
    k = np.arange(1, 2000 if n == 3 else 1500, dtype=np.float64)
    assert np.array_equal(cm_eval(f"root{n}", k ** n), k)
    sp = cm_eval(f"root{n}", np.array([0.0, np.inf, np.nan, -1.0, -np.inf]))
    assert sp[0] == 0.0 and sp[1] == np.inf and np.isnan(sp[2]) and np.isnan(sp[3]) and np.isnan(sp[4])
    # float: the double root rounded once
    xf = (10 ** rng.uniform(-37, 38, 20000)).astype(np.float32).astype(np.float64)
    gotf = cm_eval(f"root{n}", xf, precision="single")
    assert ulp_err32(gotf, root_reference(xf, n)).max() <= 1.0


@pytest.mark.parametrize("precision", ["double", "single"])
def test_min_step_is_sixteen_spacings(precision):
    """cm::min_step(t, 1.1 t1), the controller's minimum step, equals the
    formula 16 |fabs(nextafter(t, 1.1 t1)) - t| it replaces, bit for bit, over
    normal, subnormal, zero, negative, extreme and non-finite times and both
    directions (numpy's nextafter is the independent side)."""
    rng = np.random.default_rng(77)
    ft = np.float64 if precision == "double" else np.float32
    fi = np.finfo(ft)
    mag = 10 ** rng.uniform(-300 if ft is np.float64 else -37, 300 if ft is np.float64 else 37, 50000)
    t = np.concatenate([mag, -mag[:5000], rng.uniform(0, 1000, 20000),
                        2.0 ** rng.integers(-200, 100, 5000).astype(np.float64),
                        [0.0, -0.0, float(fi.tiny), float(fi.tiny) / 4, float(fi.max), -float(fi.max),
                         float(fi.smallest_subnormal), np.inf, -np.inf, np.nan]]).astype(ft)
    t1 = np.concatenate([t[:-10] * ft(1) + np.abs(rng.normal(0, 1, t.size - 10)).astype(ft) * np.abs(t[:-10]),
                         rng.uniform(-10, 10, 10).astype(ft)]).astype(ft)
    t1[::7] = t[::7] * ft(0.5)  # toward a smaller value on a seventh of the cases
    toward = (ft(1.1) * t1).astype(ft)
    with np.errstate(all="ignore"):
        want = (ft(16) * np.abs(np.abs(np.nextafter(t, toward)) - t)).astype(ft)
        got = cm_eval("min_step", t.astype(np.float64), toward.astype(np.float64), precision=precision)
    same = (got.astype(ft) == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), (t[~same][:5], toward[~same][:5], got[~same][:5], want[~same][:5])
    # the fast branch covers the usual case: a positive time short of t1
    usual = (t > fi.tiny) & (t < toward) & np.isfinite(t)
    assert usual.mean() > 0.5


@pytest.mark.parametrize("precision", ["double", "single"])
def test_root2_is_the_square_root(precision):
    """cm::root<2>, the he12 controller's (rtol/err)^(1/2): pow(x, 0.5) with
    its exponent exact is the correctly rounded square root, so the root is
    numpy's sqrt bit for bit (+0 for -0, NaN below zero, inf for inf)."""
    rng = np.random.default_rng(9)
    hi = 300 if precision == "double" else 37
    x = np.concatenate([10 ** rng.uniform(-hi, hi, 100000), rng.uniform(0, 4, 10000), [0.0, np.inf]])
    if precision == "single":
        x = x.astype(np.float32).astype(np.float64)
    got = cm_eval("root2", x, precision=precision)
    want = np.sqrt(x) if precision == "double" else np.sqrt(x.astype(np.float32)).astype(np.float64)
    assert np.array_equal(got, want)
    sp = cm_eval("root2", np.array([-0.0, -1.0, np.nan]), precision=precision)
    assert sp[0] == 0.0 and not np.signbit(sp[0]) and np.isnan(sp[1]) and np.isnan(sp[2])
