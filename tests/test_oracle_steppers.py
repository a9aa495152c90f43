"""Pins the oracle's steppers with closed forms and convergence orders.

A dropped tableau term, a wrong stage time or a transposed coefficient
changes the observed order or the closed-form error by orders of magnitude,
which these checks catch:
  * decay x' = -x and the harmonic oscillator with fixed steps: global error
    ratios on halving dt give orders 1 (euler), 2 (heun), 4 (rk4);
  * Euler on decay equals (1 - h)^n to rounding;
  * adaptive bs23/dopri5: error against the closed form shrinks with rtol
    (proportional control); the last step ends one hmin past the span end;
  * the falling body with linear drag (ORNL problem A1, the reference's
    test/test_ornl_thompson_a1.py) along the stored trajectory to 5
    significant digits, in single precision as the reference runs it.
"""
import math

import numpy as np
import pytest

from oracle import OracleEnsemble


def run_decay(stepper, dt, t_end=1.0, prec="double"):
    o = OracleEnsemble("decay", 1, 1, stepper=stepper, precision=prec, kind="base", threads=1)
    o.set_solver_params(dt, dt, 1e-12, 1e-10, 10 ** 8, 1000, 1)
    o.set_problem_data(np.ones((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, t_end))
    o.transient()
    return o.get("xf")[0], o.get("tf")[0]


@pytest.mark.parametrize("stepper,order", [("euler", 1), ("heun", 2), ("rk4", 4)])
def test_fixed_step_orders(stepper, order):
    # the loop runs while t <= t_end, so the final time overshoots by < dt:
    # compare against exp(-tf) at the actual final time
    errs = []
    for dt in (1 / 64, 1 / 128, 1 / 256):
        x, tf = run_decay(stepper, dt)
        errs.append(abs(x - math.exp(-tf)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert abs(math.log2(r1) - order) < 0.15 and abs(math.log2(r2) - order) < 0.15, (errs, r1, r2)


def test_euler_is_repeated_multiplication():
    h = 1 / 32
    x, tf = run_decay("euler", h, t_end=1.0)
    n = round(tf / h)
    assert x == pytest.approx((1 - h) ** n, rel=1e-14)


def test_rk4_harmonic_closed_form():
    o = OracleEnsemble("harmonic", 2, 1, stepper="rk4", kind="base", threads=1)
    o.set_solver_params(0.01, 0.01, 1e-6, 1e-3, 10 ** 7, 1000, 1)
    o.set_problem_data(np.array([[1.0], [0.0]]), np.array([[1.0]]))
    o.set_tspan((0.0, 2 * math.pi))
    o.transient()
    x, y = o.get("xf")
    tf = o.get("tf")[0]
    assert abs(x - math.cos(tf)) < 1e-9 and abs(y + math.sin(tf)) < 1e-9


@pytest.mark.parametrize("stepper", ["bs23", "dopri5", "he12"])
def test_adaptive_tolerance_proportionality(stepper):
    errs = []
    for rtol in (1e-4, 1e-6, 1e-8):
        o = OracleEnsemble("decay", 1, 1, stepper=stepper, kind="base", threads=1)
        o.set_solver_params(0.01, 0.5, rtol * 1e-3, rtol, 10 ** 8, 1000, 1)
        o.set_problem_data(np.ones((1, 1)), np.ones((1, 1)))
        o.set_tspan((0.0, 5.0))
        o.transient()
        tf = o.get("tf")[0]
        errs.append(abs(o.get("xf")[0] - math.exp(-tf)) / math.exp(-tf))
    assert errs[0] > errs[1] > errs[2]
    assert errs[1] < 1e-4 and errs[2] < 1e-6


def test_adaptive_end_of_span_and_dt_carry():
    # the controller clips the last step to the end of the span and then to
    # hmin = 16 ulp(t): the loop takes one extra hmin step past t1 and the
    # stored dt is that hmin (adaptive_explicit_step.clh:68-72, transient.cl:73)
    o = OracleEnsemble("decay", 1, 1, stepper="dopri5", kind="base", threads=1)
    o.set_solver_params(0.1, 0.5, 1e-9, 1e-6, 10 ** 6, 1000, 1)
    o.set_problem_data(np.ones((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, 10.0))
    o.transient()
    tf, dt = o.get("tf")[0], o.get("dt")[0]
    assert 10.0 < tf <= 10.0 + 32 * np.spacing(10.0)
    assert dt == pytest.approx(16 * np.spacing(tf), rel=0.5)


@pytest.mark.parametrize("prec", ["single", "double"])
def test_linear_drag_trajectory_five_digits(prec):
    """Reference test/test_ornl_thompson_a1.py: rk4, dt 1e-3, t in [0, 5]."""
    m, w, k, H = 0.25, 8.0, 2.0, 10.0
    o = OracleEnsemble("linear_drag", 2, 4, 1, stepper="rk4", precision=prec, kind="trajectory", threads=1)
    o.set_solver_params(0.001, 0.001, 1e-6, 1e-4, 20000, 20000, 1)
    o.set_problem_data(np.zeros((2, 1)), np.array([[m], [w], [k], [H]]))
    o.set_tspan((0.0, H / 2.0))
    o.trajectory()
    n = int(o.get("nStored")[0]) + 1
    t = o.get("t")[:n]
    x = o.get("x").reshape(-1, 2)[:n]
    y1 = 4.0 * (t + 1.0 / 8.0 * np.exp(-8.0 * t) - 1.0 / 8.0)
    y2 = 4.0 * (1.0 - np.exp(-8.0 * t))
    for a, b in ((x[1:, 0], y1[1:]), (x[1:, 1], y2[1:])):
        np.testing.assert_allclose(a, b, rtol=5e-5, atol=1e-6 if prec == "single" else 0)
    assert n == 5001 or n == 5002


def first_adaptive_step(stepper, lam, t0, t1, max_steps=1):
    o = OracleEnsemble("decay", 1, 1, stepper=stepper, kind="base", threads=1)
    o.set_solver_params(1.0, 1.0, 1e-12, 1e-6, max_steps, 1000, 1)
    o.set_problem_data(np.ones((1, 1)), np.full((1, 1), lam))
    o.set_tspan((t0, t1))
    o.transient()
    return o.get("xf")[0], o.get("tf")[0], o.get("dt")[0]


@pytest.mark.parametrize("stepper", ["bs23", "dopri5", "he12"])
def test_controller_halves_past_one_hundred_trials(stepper):
    """x' = -1e31 x from t = 0 with dt = 1: the first trial fails by ~1e31,
    so the step shrinks to 0.1 (the maximal first shrink: 0.5 for bs23, 0.1
    for dopri5) and then halves until lambda h is small; that takes more than
    100 halvings.  The reference keeps halving until hmin (~8e-323 at t = 0,
    adaptive_explicit_step.clh:44-50), so the step is accepted at
    h = s 2^-k (exactly: t = 0 + h is exact) and x(h) = exp(-lambda h).  A
    bound on trials of 100 would instead fail the step at hmin (t stays 0)."""
    x, tf, dt = first_adaptive_step(stepper, 1e31, 0.0, 1.0)
    shrink = 0.1 if stepper == "dopri5" else 0.5
    assert tf > 0.0
    k = math.log2(shrink / tf)
    assert k == round(k) and k > 100, k
    assert tf == shrink * 2.0 ** -round(k)
    assert x == pytest.approx(math.exp(-1e31 * tf), rel=1e-5)
    assert 1e31 * tf < 1.0  # the accepted step resolves the decay


@pytest.mark.parametrize("stepper", ["bs23", "dopri5", "he12"])
def test_controller_fails_a_stalled_trial_at_hmin(stepper):
    """t0 = 1 - 13 2^-53, just below a power of two: hmin = 16 ulp = 2^-49,
    but t0 + hmin rounds (to even) to 1 + 2^-51, a step of 17 ulp > hmin.  A
    right-hand side whose error no step can meet (x' = -1e40 x) is rejected
    down to that trial, which then repeats forever in the reference's loop
    (its exit needs a step <= hmin).  Here it fails at hmin: each of the
    max_steps = 3 steps leaves t and x unchanged with dt = hmin."""
    t0 = 1.0 - 13 * 2.0 ** -53
    assert (t0 + 2.0 ** -49) - t0 == 17 * 2.0 ** -53
    x, tf, dt = first_adaptive_step(stepper, 1e40, t0, 2.0, max_steps=3)
    assert tf == t0 and x == 1.0 and dt == 2.0 ** -49


def test_he12_step_closed_form():
    """Heun-Euler 1(2) on x' = -lambda x from x = 1 with step h: the Euler
    predictor 1 - lambda h, the trapezoidal state 1 - lambda h + (lambda h)^2 / 2,
    the error estimate (lambda h)^2 / 2; the next step is h times
    min(5, 0.8 sqrt(rtol / err)) (error order 1: the square root), and a
    rejected first trial shrinks by at most 0.5 (adaptive_he12.clh)."""
    lam, h, rtol = 1.0, 0.1, 0.1
    o = OracleEnsemble("decay", 1, 1, stepper="he12", kind="base", threads=1)
    o.set_solver_params(h, 1.0, 1e-12, rtol, 1, 1000, 1)
    o.set_problem_data(np.ones((1, 1)), np.full((1, 1), lam))
    o.set_tspan((0.0, 10.0))
    o.transient()
    x, tf, dt = o.get("xf")[0], o.get("tf")[0], o.get("dt")[0]
    err = (lam * h) ** 2 / 2
    assert tf == h
    assert x == pytest.approx(1 - lam * h + (lam * h) ** 2 / 2, rel=1e-15)
    assert dt == pytest.approx(h * 0.8 * math.sqrt(rtol / err), rel=1e-15)
    # lambda h = 10: error 50 > rtol, the first shrink is max(0.5, 0.8 sqrt(2e-5)) = 0.5
    o = OracleEnsemble("decay", 1, 1, stepper="he12", kind="base", threads=1)
    o.set_solver_params(h, 1.0, 1e-12, 1e-3, 1, 1000, 1)
    o.set_problem_data(np.ones((1, 1)), np.full((1, 1), 100.0))
    o.set_tspan((0.0, 10.0))
    o.transient()
    tf = o.get("tf")[0]
    # rejected at 0.1, then 0.05, then halved until 100 h (1 - ...) meets rtol
    k = math.log2(0.05 / tf)
    assert k == round(k) and tf <= 0.05
    lh = 100.0 * tf
    assert lh * lh / 2 <= 1e-3 < (2 * lh) ** 2 / 2
