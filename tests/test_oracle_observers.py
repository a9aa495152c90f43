"""Pins the oracle's observers with signals whose features are known in
closed form, and with the reference's own observer tests:
  * test/test_features.py: x' = cos t from 0 over [0, 4 pi] with thresh2:
    up-crossings at pi/4 and 9 pi/4, down-crossings at 3 pi/2 and 7 pi/2,
    event count exactly 2;
  * test/test_aux_values.py: aux constants 1 and -2 reported exactly, the
    mean/min/max of x + 1 ~ 1/0/2;
  * test/test_vdp.py: Van der Pol period against its asymptotic formula;
  * harmonic oscillator: period 2 pi / omega and amplitude for localmax,
    thresh2 and nhood2.
"""
import math

import numpy as np
import pytest

from oracle import OracleEnsemble

OBS_BASE = dict(e_var_ix=0, f_var_ix=0, max_event_count=100, max_event_timestamps=0, min_amp=0.0, min_imi=0.0,
                nhood_radius=0.05, x_up_threshold=0.3, x_down_threshold=0.2, dx_up_threshold=0.0,
                dx_down_threshold=0.0, eps_dx=0.0)


def feature_sim(model, nv, npar, naux, stepper, observer, prec, sp, obs=None):
    op = dict(OBS_BASE)
    op.update(obs or {})
    o = OracleEnsemble(model, nv, npar, naux, stepper=stepper, observer=observer, precision=prec,
                       n_store=op["max_event_timestamps"], threads=2)
    o.set_solver_params(*sp)
    o.set_observer_params(**op)
    return o


def names_thresh2(vars_, aux, S):
    f = [f"{p} {s}" for s in ("period", "peaks", "upDuration", "downDuration", "duty", "activeDip")
         for p in ("max", "min", "mean")]
    for v in vars_:
        f += [f"max {v}", f"min {v}", f"mean {v}", f"max d{v}/dt", f"min d{v}/dt"]
    for a in aux:
        f += [f"max {a}", f"min {a}", f"mean {a}"]
    for j in range(S):
        f += [f"up event time {j}", f"down event time {j}"]
    return f + ["event count", "step count", "max dt", "min dt", "mean dt"]


@pytest.mark.parametrize("prec", ["single", "double"])
def test_thresh2_sine_event_times(prec):
    o = feature_sim("sine", 1, 1, 3, "rk4", "thresh2", prec, (0.001, 0.001, 1e-6, 1e-3, 10 ** 7, 10 ** 7, 1),
                    dict(min_amp=0.5, x_up_threshold=(2 + math.sqrt(2)) / 4, x_down_threshold=0.001,
                         dx_up_threshold=0.001, dx_down_threshold=0.001, max_event_timestamps=3))
    o.set_problem_data(np.zeros((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, 4 * math.pi))
    o.features()
    F = dict(zip(names_thresh2(["x"], ["xp1", "pos", "neg"], 3), o.get("F")))
    assert F["event count"] == 2
    assert abs(F["up event time 0"] - math.pi / 4) < 1e-2
    assert abs(F["up event time 1"] - 9 * math.pi / 4) < 1e-2
    assert abs(F["down event time 0"] - 3 * math.pi / 2) < 1e-2
    assert abs(F["down event time 1"] - 7 * math.pi / 2) < 1e-2
    assert F["up event time 2"] == 0 and F["down event time 2"] == 0


@pytest.mark.parametrize("prec", ["single", "double"])
def test_thresh2_aux_values(prec):
    o = feature_sim("sine", 1, 1, 3, "rk4", "thresh2", prec, (0.1, 0.1, 1e-6, 1e-3, 10 ** 7, 10 ** 7, 1))
    o.set_problem_data(np.zeros((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, 400 * math.pi))
    o.features()
    F = dict(zip(names_thresh2(["x"], ["xp1", "pos", "neg"], 0), o.get("F")))
    assert F["mean xp1"] == pytest.approx(1.0, abs=1e-2)
    assert F["min xp1"] == pytest.approx(0.0, abs=1e-2)
    assert F["max xp1"] == pytest.approx(2.0, abs=1e-2)
    assert F["mean pos"] == 1.0 and F["min pos"] == 1.0 and F["max pos"] == 1.0
    assert F["min neg"] == -2.0 and F["max neg"] == -2.0


def vdp_period(mu):
    if mu < 0:
        return 0.0
    if mu < 2:
        return 2 * math.pi * (1 + mu ** 2 / 16)
    return min(2 * math.pi * (1 + mu ** 2 / 16), (3 - 2 * math.log(2)) * mu + 3 * 2.2338 / mu ** (1 / 3))


@pytest.mark.parametrize("prec", ["single", "double"])
def test_thresh2_van_der_pol_period(prec):
    mus = np.array([-1, 0, 0.01, 0.1, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 5, 6])
    o = feature_sim("van_der_pol", 2, 1, 0, "dopri5", "thresh2", prec, (0.1, 1.0, 1e-6, 1e-3, 20000, 20000, 1))
    o.set_problem_data(np.ones((2, mus.size)), mus[None, :])
    o.set_tspan((0.0, 1000.0))
    o.features()
    F = o.get("F").reshape(-1, mus.size)
    for mu, p in zip(mus, F[0]):  # "max period"
        assert p == pytest.approx(vdp_period(mu), rel=1e-2, abs=1.0), mu


def harmonic_features(observer, omega=2.0, A=1.5, prec="double", obs=None):
    o = feature_sim("harmonic", 2, 1, 0, "rk4", observer, prec, (0.001, 0.001, 1e-6, 1e-3, 10 ** 7, 10 ** 7, 1), obs)
    o.set_problem_data(np.array([[A], [0.0]]), np.array([[omega]]))
    o.set_tspan((0.0, 40.0))
    o.features()
    return o.get("F")


def test_localmax_harmonic_period_and_amplitude():
    omega, A = 2.0, 1.5
    F = harmonic_features("localmax", omega, A)
    # IMI max/min/mean, amplitude max/min/mean
    assert F[2] == pytest.approx(2 * math.pi / omega, abs=2e-3)
    assert F[5] == pytest.approx(2 * A, abs=1e-4)
    # per variable: max, min, mean, max dx, min dx (x then y)
    assert F[6] == pytest.approx(A, abs=1e-6) and F[7] == pytest.approx(-A, abs=1e-4)
    assert F[9] == pytest.approx(omega * A, abs=1e-4)  # max dx/dt = omega*A
    # event count: maxima of A cos(2 t) at t = pi k, k = 1..12 in (0, 40]
    assert F[-2] == 12


def test_thresh2_harmonic_period():
    omega = 2.0
    F = harmonic_features("thresh2", omega)
    assert F[2] == pytest.approx(2 * math.pi / omega, abs=2e-3)   # mean period
    assert F[5] == 1.0                                          # one peak per period
    # up: rising crossing of -0.4 A (phase 2 pi - acos(-0.4)); down: falling crossing of
    # -0.6 A (phase acos(-0.6)); duty = up-duration / period
    duty = (math.acos(-0.6) + math.acos(-0.4)) / (2 * math.pi)
    assert F[14] == pytest.approx(duty, abs=5e-3)  # mean duty


def test_nhood2_harmonic_period():
    omega = 2.0
    F = harmonic_features("nhood2", omega)
    assert F[2] == pytest.approx(2 * math.pi / omega, abs=2e-3)
    # nhood centre: first downward crossing of min + 0.2 range in x
    nv_block = F[6:6 + 7]
    assert nv_block[3] == pytest.approx(3.0, abs=1e-5)          # range x = 2A
    assert nv_block[4] == pytest.approx(-1.5 + 0.2 * 3.0, abs=1e-2)


def test_localmax_no_events_for_monotone_decay():
    o = feature_sim("decay", 1, 1, 0, "rk4", "localmax", "double", (0.01, 0.01, 1e-6, 1e-3, 10 ** 6, 10 ** 6, 1),
                    dict(max_event_timestamps=2))
    o.set_problem_data(np.ones((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, 5.0))
    o.features()
    F = o.get("F")
    assert F[-2] == 0 and np.all(F[:6] == 0)
    assert F[-1] == 501 or F[-1] == 502  # step count (t <= 5 loop)


def test_basicall_harmonic_closed_forms():
    """x = A cos(w t), y = -A sin(w t) on [0, T]: extents +-A, slopes +-wA, and
    time means A sin(wT)/(wT) and -A (1 - cos(wT))/(wT)."""
    omega, A, T = 2.0, 1.5, 40.0
    F = harmonic_features("basicall", omega, A)
    assert len(F) == 5 * 2 + 1
    assert F[0] == pytest.approx(A, abs=1e-6) and F[1] == pytest.approx(-A, abs=1e-5)
    assert F[2] == pytest.approx(A * math.sin(omega * T) / (omega * T), abs=1e-4)
    assert F[3] == pytest.approx(omega * A, abs=1e-5) and F[4] == pytest.approx(-omega * A, abs=1e-5)
    assert F[5] == pytest.approx(A, abs=1e-5) and F[6] == pytest.approx(-A, abs=1e-5)
    assert F[7] == pytest.approx(-A * (1 - math.cos(omega * T)) / (omega * T), abs=1e-4)
    assert F[-1] in (40000, 40001, 40002)  # step count of the t <= T loop with dt 1e-3


def test_basic_reports_the_feature_variable():
    """basic with f_var_ix = 1 gives exactly basicall's block for y."""
    Fb = harmonic_features("basic", obs=dict(f_var_ix=1))
    Fa = harmonic_features("basicall")
    assert len(Fb) == 6
    assert np.array_equal(Fb[:5], Fa[5:10]) and Fb[5] == Fa[-1]


def test_nhood1_harmonic_returns():
    """Centre = state at the first minimum of x (t = pi/2); an event on every
    entry into the ball: one just after the centre is found, then one per
    return, entering 0.1 rad of phase (radius 0.05 of the range-normalised
    circle of radius 0.5) before t = pi/2 + k pi, k = 1..12."""
    omega = 2.0
    F = harmonic_features("nhood1", omega)
    names_block = 6 + 5 * 2
    period_count, step_count = F[names_block], F[names_block + 1]
    assert period_count == 12
    assert F[0] == pytest.approx(math.pi, abs=2e-3)          # max period
    assert F[1] == pytest.approx(math.pi - 0.05, abs=5e-3)   # first period: from the centre to the first entry
    assert F[5] == pytest.approx(1.0)                        # one maximum of x per return
    assert F[6] == pytest.approx(1.5, abs=1e-6)              # max x
    assert step_count in (40000, 40001, 40002)
    # step-count mean of x (uniform steps): A sin(wT)/(wT)
    assert F[8] == pytest.approx(1.5 * math.sin(80.0) / 80.0, abs=1e-4)


def test_nhood1_without_events_reports_wrapped_period_count():
    o = feature_sim("decay", 1, 1, 0, "rk4", "nhood1", "double", (0.01, 0.01, 1e-6, 1e-3, 10 ** 6, 10 ** 6, 1))
    o.set_problem_data(np.ones((1, 1)), np.ones((1, 1)))
    o.set_tspan((0.0, 5.0))
    o.features()
    F = o.get("F")
    assert np.all(F[:6] == 0)
    assert F[6 + 5] == 2.0 ** 32 - 1  # eventcount - 1 as unsigned with no event


def test_thresh2_harmonic_durations_and_active_dip():
    """x = A cos(w t) with the default thresholds: the up event is the rising
    crossing of -0.4 A (phase 2 pi - acos(-0.4)), the down transition the
    falling crossing of -0.6 A (phase acos(-0.6)).  Up duration
    (acos(-0.6) + acos(-0.4)) / w, down duration the rest of the period, duty
    their ratio to the period.  activeDip = time mean of x over the down state
    (phases acos(-0.6) .. 2 pi - acos(-0.4)) minus the last local minimum
    (-A): A (sin th2 - sin th1) / (th2 - th1) + A.  Events are step points
    (dt = 1e-3), so durations are exact to a step."""
    omega, A = 2.0, 1.5
    F = dict(zip(names_thresh2(["x", "y"], [], 0), harmonic_features("thresh2", omega, A)))
    th1, th2 = math.acos(-0.6), 2 * math.pi - math.acos(-0.4)
    up = (math.acos(-0.6) + math.acos(-0.4)) / omega
    down = 2 * math.pi / omega - up
    for stat in ("max", "min", "mean"):
        assert F[f"{stat} upDuration"] == pytest.approx(up, abs=2.5e-3), stat
        assert F[f"{stat} downDuration"] == pytest.approx(down, abs=2.5e-3), stat
        assert F[f"{stat} duty"] == pytest.approx(up / (up + down), abs=2e-3), stat
    dip = A * (math.sin(th2) - math.sin(th1)) / (th2 - th1) + A
    assert dip == pytest.approx(0.26603, abs=1e-5)  # 0.17735 A
    for stat in ("max", "min", "mean"):
        assert F[f"{stat} activeDip"] == pytest.approx(dip, abs=5e-3), stat
    # a mistaken dip (e.g. the down-state mean alone, or the mean over the
    # whole period, 0) is off by at least 0.26
    assert F["event count"] == 13  # up crossings at t = 2.150 + pi k <= 40


def two_tone_features(observer, b=0.6, obs=None):
    o = feature_sim("two_tone", 5, 1, 0, "rk4", observer, "double",
                    (1.0 / 256, 1.0, 1e-6, 1e-3, 10 ** 7, 10 ** 7, 1),
                    dict(dict(e_var_ix=4, f_var_ix=4), **(obs or {})))
    o.set_problem_data(np.array([[1.0], [0.0], [1.0], [0.0], [1.0 + b]]), np.array([[b]]))
    o.set_tspan((0.0, 20 * math.pi))
    o.features()
    return o.get("F")


def test_thresh2_and_nhood2_count_two_maxima_per_period():
    """s = cos t + 0.6 cos 2t: two local maxima per period (1.6 at t = 0,
    -0.4 at t = pi).  The thresholds (0.3 / 0.2 of the range above the
    minimum -0.808) put one up event per period, so thresh2 counts 2 peaks
    per period; nhood2's centre is the first downward crossing of 0.2 of the
    range and one exit from its ball happens per period, with 2 maxima in
    between.  A rule that counted sign changes of ds differently (e.g. every
    extremum, or only the largest maximum) gives 4 or 1."""
    F = dict(zip(names_thresh2(["u1", "v1", "u2", "v2", "s"], [], 0), two_tone_features("thresh2")))
    assert F["mean period"] == pytest.approx(2 * math.pi, abs=1e-2)
    assert F["max peaks"] == 2 and F["min peaks"] == 2 and F["mean peaks"] == 2
    assert F["max s"] == pytest.approx(1.6, abs=1e-6)  # sampled at steps of 1/256
    smin = -1 / 2.4 + 0.6 * (2 * (1 / 2.4) ** 2 - 1)
    assert F["min s"] == pytest.approx(smin, abs=1e-4)
    Fn = two_tone_features("nhood2")
    # nhood2: period max/min/mean, then peaks max/min/mean
    assert Fn[2] == pytest.approx(2 * math.pi, abs=1e-2)
    assert Fn[3] == 2 and Fn[4] == 2 and Fn[5] == 2


def test_localmax_alternating_maxima():
    """localmax on the same s: every maximum is an event, so inter-maximum
    intervals are pi; amplitudes (maximum minus the last local minimum)
    alternate between 1.6 - smin and -0.4 - smin."""
    smin = -1 / 2.4 + 0.6 * (2 * (1 / 2.4) ** 2 - 1)
    F = two_tone_features("localmax")
    assert F[0] == pytest.approx(math.pi, abs=1e-2) and F[1] == pytest.approx(math.pi, abs=1e-2)
    assert F[3] == pytest.approx(1.6 - smin, abs=1e-4)   # max amplitude
    assert F[4] == pytest.approx(-0.4 - smin, abs=1e-4)  # min amplitude
