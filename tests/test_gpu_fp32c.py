"""fp32-compensated mode against the FP64 CPU oracle.

Stated tolerances (adaptive FP32 states within a stated rtol; each feature's
error against the FP64 oracle stated):
  * fixed-step rk4 Van der Pol over 2,000 steps: |x - x_fp64| <= 2e-5 (1 + |x_fp64|);
  * fixed-step Euler on decay over 1e5 steps: the Kahan-compensated state
    error against FP64 is at least 10x smaller than plain single precision;
  * adaptive Heun-Euler on Van der Pol: the state within 3x FP64's own
    truncation error (rtol against rtol / 100) of the FP64 oracle on >= 95 %
    of the members, median below 1e-4 (1 + |x|);
  * dopri5 lactotroph features on a parameter sweep that crosses regime
    boundaries, over members whose event counts agree (>= 90 %): period,
    peaks, min/mean V, mean c median relative error <= 1e-4, 95th percentile
    <= 1e-3; sampled spike peak max V median <= 5e-4, 95th percentile <= 1e-2.
"""
import numpy as np
import pytest

from helpers import Pair

pytestmark = pytest.mark.gpu


def test_rk4_vdp_state_close_to_fp64(gpu_available):
    n = 129
    mu = np.linspace(0.1, 3.0, n)
    x0 = np.vstack([2.0 * np.ones(n), np.zeros(n)])
    c = Pair("van_der_pol", "rk4", "", "fp32c", kind="base", sp=(0.01, 1.0, 1e-6, 1e-3, 10 ** 7, 10, 1),
             oracle=False)
    c.set_problem_data(x0, mu[None, :])
    c.set_tspan(0.0, 20.0)
    c.transient()
    ref = Pair("van_der_pol", "rk4", "", "double", kind="base", sp=(0.01, 1.0, 1e-6, 1e-3, 10 ** 7, 10, 1))
    ref.set_problem_data(x0, mu[None, :])
    ref.set_tspan(0.0, 20.0)
    ref.orc.transient()
    a = np.array(c.gpu.get_xf())
    b = ref.orc.get("xf")
    err = float(np.max(np.abs(a - b) / (1.0 + np.abs(b))))
    assert err < 2e-5, err


def test_kahan_beats_plain_single(gpu_available):
    n = 64
    lam = np.linspace(0.1, 0.5, n)
    sp = (1e-4, 1e-4, 1e-6, 1e-3, 10 ** 7, 10, 1)
    errs = {}
    ref = Pair("decay", "euler", "", "double", kind="base", sp=sp)
    ref.set_problem_data(np.ones((1, n)), lam[None, :])
    ref.set_tspan(0.0, 10.0)
    ref.orc.transient()
    b = ref.orc.get("xf")
    for prec in ("single", "fp32c"):
        p = Pair("decay", "euler", "", prec, kind="base", sp=sp, oracle=False)
        p.set_problem_data(np.ones((1, n)), lam[None, :])
        p.set_tspan(0.0, 10.0)
        p.transient()
        errs[prec] = float(np.max(np.abs(np.array(p.gpu.get_xf()) - b) / b))
    assert errs["fp32c"] * 10 < errs["single"], errs


def test_dopri5_lactotroph_features_vs_fp64(gpu_available):
    n = 256
    g = np.linspace(0.0, 10.0, n)
    b = np.linspace(0.0, 2.0, n)[::-1]
    x0 = np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n))
    pars = np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), b, g])
    sp = (0.1, 1.0, 1e-8, 1e-6, 10 ** 8, 10, 1)
    c = Pair("lactotroph", "dopri5", "thresh2", "fp32c", sp=sp, oracle=False)
    c.set_problem_data(x0, pars)
    c.set_tspan(0.0, 1000.0)
    c.transient()
    c.shift_x0()
    c.set_tspan(0.0, 4000.0)
    c.features()
    ref = Pair("lactotroph", "dopri5", "thresh2", "double", sp=sp)
    ref.set_problem_data(x0, pars)
    ref.set_tspan(0.0, 1000.0)
    ref.orc.transient()
    ref.orc.shift_x0()
    ref.set_tspan(0.0, 4000.0)
    ref.orc.features()
    names = c.gpu.get_feature_names()
    A = np.array(c.gpu.get_f()).reshape(-1, n)
    B = ref.orc.get("F").reshape(-1, n)
    ec = names.index("event count")
    same_events = A[ec] == B[ec]
    assert same_events.mean() > 0.9
    # median / 95th-percentile bounds; the sampled spike peak (max v) depends on
    # where the adaptive steps land around the maximum, so its bound is wider
    tol = {"mean period": (1e-4, 1e-3), "mean peaks": (1e-4, 1e-3), "min v": (1e-4, 1e-3),
           "mean v": (1e-4, 1e-3), "mean c": (1e-4, 1e-3), "max v": (5e-4, 1e-2)}
    for key, (med, p95) in tol.items():
        j = names.index(key)
        ok = same_events & (B[j] != 0)
        rel = np.abs(A[j][ok] - B[j][ok]) / np.abs(B[j][ok])
        assert float(np.median(rel)) <= med, (key, float(np.median(rel)))
        assert float(np.percentile(rel, 95)) <= p95, (key, float(np.percentile(rel, 95)))


def test_he12_vdp_state_within_truncation_of_fp64(gpu_available):
    """Heun-Euler 1(2) in fp32c: the Van der Pol state after t = 20 differs
    from FP64 at the same tolerance by no more than 3x FP64's own truncation
    error (FP64 at rtol against rtol / 100) plus 1e-5 (1 + |x|), on at least
    95 % of the members, with a median difference below 1e-4 (1 + |x|); the
    two take different step sequences, so their difference is of the
    method's truncation size, not of float rounding."""
    n = 129
    mu = np.linspace(0.1, 3.0, n)
    x0 = np.vstack([2.0 * np.ones(n), np.zeros(n)])
    sp = (0.01, 1.0, 1e-7, 1e-5, 10 ** 8, 10, 1)
    c = Pair("van_der_pol", "he12", "", "fp32c", kind="base", sp=sp, oracle=False)
    c.set_problem_data(x0, mu[None, :])
    c.set_tspan(0.0, 20.0)
    c.transient()
    refs = []
    for scale in (1.0, 1e-2):
        r = Pair("van_der_pol", "he12", "", "double", kind="base",
                 sp=(sp[0], sp[1], sp[2] * scale, sp[3] * scale) + sp[4:])
        r.set_problem_data(x0, mu[None, :])
        r.set_tspan(0.0, 20.0)
        r.orc.transient()
        refs.append(r.orc.get("xf"))
    a = np.array(c.gpu.get_xf())
    d, t = np.abs(a - refs[0]), np.abs(refs[0] - refs[1])
    within = d <= 3 * t + 1e-5 * (1 + np.abs(refs[0]))
    assert np.mean(within) >= 0.95, (float(np.mean(within)), float(np.max(d)))
    assert float(np.median(d / (1 + np.abs(refs[0])))) < 1e-4
