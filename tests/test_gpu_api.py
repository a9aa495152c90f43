"""The reference's own Python tests, run through clode_amd's Python API on the
GPU (test/test_vdp.py, test/test_features.py, test/test_aux_values.py,
test/test_ornl_thompson_a1.py), plus ensemble construction semantics of
clode/solver.py:254-462."""
import math
import os
from typing import List

import numpy as np
import pytest

import clode_amd as clode

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "clode_amd", "models")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vdp_period(mu):
    if mu < 0:
        return 0.0
    if mu < 2:
        return 2 * math.pi * (1 + mu ** 2 / 16)
    return min(2 * math.pi * (1 + mu ** 2 / 16), (3 - 2 * math.log(2)) * mu + 3 * 2.2338 / mu ** (1 / 3))


def vdp_rhs(t: float, var: List[float], par: List[float], derivatives: List[float], aux: List[float],
            wiener: List[float]) -> None:
    mu: float = par[0]
    x: float = var[0]
    y: float = var[1]
    dx: float = y
    dy: float = mu * (1 - x * x) * y - x
    derivatives[0] = dx
    derivatives[1] = dy


@pytest.mark.parametrize("source", ["cl", "xpp", "python"])
@pytest.mark.parametrize("precision", [True, False, "fp32c"])
def test_vdp_dormand_prince(gpu_available, precision, source, tmp_path):
    """test/test_vdp.py:47-102 with the model file, the reference's XPP test
    file (test/test_xpp_parser.py:19) and a Python right-hand side."""
    if source == "cl":
        kw = dict(src_file=os.path.join(MODELS, "van_der_pol.cl"))
    elif source == "xpp":
        xpp = tmp_path / "van_der_pol_oscillator.xpp"
        with open(os.path.join(GOLDEN, "xpp", "van_der_pol_oscillator.xpp")) as f:
            xpp.write_text(f.read())
        kw = dict(src_file=str(xpp))
    else:
        kw = dict(rhs_equation=vdp_rhs)
    sim = clode.FeatureSimulator(**kw, variables={"x": 1.0, "y": 1.0},
                                 parameters={"mu": 1.0}, observer=clode.Observer.threshold_2,
                                 stepper=clode.Stepper.dormand_prince, t_span=(0.0, 1000.0), max_store=20000,
                                 max_steps=20000, single_precision=precision)
    mus = [-1, 0, 0.01, 0.1, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0] + list(range(5, 7))
    sim.set_ensemble(parameters={"mu": mus})
    sim.features()
    periods = sim.get_observer_results().get_var_max("period")
    for i, mu in enumerate(mus):
        assert np.isclose(periods[i], vdp_period(mu), rtol=0.01, atol=1), (mu, periods[i])


def test_sine_curve_timestamps(gpu_available):
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "sine.cl"), variables={"x": 0.0},
                                 parameters={"omega": 1.0}, aux=["xp1", "pos", "neg"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.rk4, dtmax=0.001,
                                 dt=0.001, t_span=(0.0, 4 * math.pi), event_var="x", feature_var="x",
                                 observer_min_x_amp=0.5, observer_x_up_thresh=(2 + math.sqrt(2)) / 4,
                                 observer_x_down_thresh=0.001, observer_dx_down_thresh=0.001,
                                 observer_dx_up_thresh=0.001, observer_max_event_count=100,
                                 observer_max_event_timestamps=3)
    out = sim.features()
    assert int(out.get_var_count("event")) == 2
    up, down = out.get_event_data("up"), out.get_event_data("down")
    assert len(up) == 2 and len(down) == 2
    assert np.isclose(up[0], math.pi / 4, atol=0.01) and np.isclose(up[1], 9 * math.pi / 4, atol=0.01)
    assert np.isclose(down[0], 3 * math.pi / 2, atol=0.01) and np.isclose(down[1], 7 * math.pi / 2, atol=0.01)
    assert list(out.get_timestamps("up")) == list(up)


def test_aux_values(gpu_available):
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "sine.cl"), variables={"x": 0.0},
                                 parameters={"omega": 1.0}, aux=["xp1", "pos", "neg"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.rk4, dtmax=0.1, dt=0.1,
                                 t_span=(0.0, 400 * math.pi), event_var="x", feature_var="x")
    out = sim.features()
    assert out.get_var_mean("xp1") == pytest.approx(1.0, abs=1e-2)
    assert out.get_var_min("xp1") == pytest.approx(0.0, abs=1e-2)
    assert out.get_var_max("xp1") == pytest.approx(2.0, abs=1e-2)
    assert out.get_var_mean("pos") == 1.0 and out.get_var_min("pos") == 1.0 and out.get_var_max("pos") == 1.0
    assert out.get_var_min("neg") == -2.0 and out.get_var_max("neg") == -2.0


def test_ornl_thompson_a1(gpu_available):
    m, w, k, H = 1 / 4, 8.0, 2.0, 10.0
    sim = clode.TrajectorySimulator(src_file=os.path.join(MODELS, "linear_drag.cl"), variables={"y1": 0.0, "y2": 0.0},
                                    parameters={"m": m, "w": w, "k": k, "H": H}, aux=["g1"], num_noise=0, dt=0.001,
                                    dtmax=0.001, stepper=clode.Stepper.rk4, t_span=(0.0, H / 2.0), max_store=20000,
                                    max_steps=20000)
    traj = sim.trajectory()
    for tt, y1, y2 in zip(traj.t[1:], traj.x["y1"][1:], traj.x["y2"][1:]):
        np.testing.assert_approx_equal(y1, 4.0 * (tt + 1.0 / 8.0 * math.exp(-8.0 * tt) - 1.0 / 8.0), significant=5)
        np.testing.assert_approx_equal(y2, 4.0 * (1.0 - math.exp(-8.0 * tt)), significant=5)
    assert np.allclose(traj.aux["g1"], traj.x["y1"] - H)


def test_ensemble_construction_semantics(gpu_available):
    sim = clode.Simulator(src_file=os.path.join(MODELS, "lorenz.cl"), variables={"x": 1.0, "y": 2.0, "z": 3.0},
                          parameters={"sigma": 10.0, "rho": 28.0, "beta": 8 / 3}, stepper=clode.Stepper.rk4,
                          dt=0.001, t_span=(0.0, 1.0), single_precision=False)
    sim.set_ensemble(parameters={"rho": np.linspace(0, 30, 12).reshape(3, 4)})
    assert sim._ensemble_size == 12 and sim._ensemble_shape == (3, 4)
    x0 = sim.get_initial_state()
    assert x0.shape == (12, 3) and np.all(x0 == [1.0, 2.0, 3.0])
    with pytest.raises(ValueError):
        sim.set_ensemble(parameters={"nope": 1.0})
    with pytest.raises(ValueError):
        sim.set_ensemble(variables={"x": [1, 2]}, parameters={"rho": [1, 2, 3]})
    with pytest.raises(ValueError):
        sim.set_ensemble()
    xf = sim.transient(fetch_results=True)
    assert xf.shape == (12, 3) and np.all(np.isfinite(xf))
    assert np.array_equal(sim.get_initial_state(), xf)  # update_x0 shifts xf -> x0
    assert sim.get_final_time().shape == (3, 4)
    sim.set_repeat_ensemble(5)
    assert sim.get_initial_state().shape == (5, 3)
    sim.seed_rng(1)
    sim.set_tspan((0.0, 0.5))
    assert sim.get_tspan() == (0.0, 0.5)
    sim.shift_tspan()
    assert sim.get_tspan() == (0.5, 1.0)
    assert "rk4" in sim.get_available_steppers()


def harmonic_rhs(t: float, vars: List[float], p: List[float], dy: List[float], aux: List[float],
                 w: List[float]) -> None:
    k: float = p[0]
    x: float = vars[0]
    y: float = vars[1]
    dy[0] = y
    dy[1] = -k * x


@pytest.mark.parametrize("observer", list(clode.Observer))
def test_observers_on_the_unit_circle(gpu_available, observer):
    """The harmonic-oscillator cases of test/test_observers.py:47-138 (a
    script there; assertions here): x = -cos t, y = sin t from (-1, 0), rk4
    dt 0.01 over 10 periods, feature variable y."""
    sim = clode.FeatureSimulator(rhs_equation=harmonic_rhs, variables={"x": -1.0, "y": 0.0}, parameters={"k": 1.0},
                                 dt=0.01, t_span=(0.0, 10 * 2 * math.pi), observer=observer, event_var="y",
                                 feature_var="y", observer_min_x_amp=0.0, observer_x_up_thresh=0.3,
                                 observer_x_down_thresh=0.2, observer_neighbourhood_radius=0.05,
                                 observer_max_event_timestamps=10, single_precision=False)
    out = sim.features()
    names = out.get_feature_names()
    state = sim.get_final_state()[0]
    # back at the start after 10 periods (the last step ends up to dt past 20 pi)
    assert np.allclose(state, [-1.0, 0.0], atol=0.02)
    if "max x" in names:
        assert out.get_var_max("x") == pytest.approx(1.0, abs=1e-6)
        assert out.get_var_min("x") == pytest.approx(-1.0, abs=1e-6)
    if "mean x" in names:
        assert out.get_var_mean("x") == pytest.approx(0.0, abs=2e-3)
        assert out.get_var_mean("y") == pytest.approx(0.0, abs=2e-3)
    if "mean amplitude" in names:
        assert out.get_var_mean("amplitude") == pytest.approx(2.0, abs=1e-3)
    if "mean IMI" in names:
        for stat in (out.get_var_max, out.get_var_min, out.get_var_mean):
            assert stat("IMI") == pytest.approx(2 * math.pi, abs=0.02)
    if "mean period" in names:
        assert out.get_var_mean("period") == pytest.approx(2 * math.pi, abs=0.02)
    if "event count" in names:
        # one event per period (the neighbourhood observers start counting at
        # their trigger point, so they see one fewer)
        assert int(out.get_var_count("event")) in (9, 10)
    if "step count" in names:
        assert int(out.get_var_count("step")) == int(10 * 2 * math.pi / 0.01) + 1


def test_logger_and_device_printing(gpu_available, capfd):
    """test/test_logger.py:4-28: device printing follows the log level."""
    sim = clode.TrajectorySimulator(src_file=os.path.join(MODELS, "van_der_pol.cl"), variables={"x": 0.0, "y": 1.0},
                                    parameters={"mu": 1.0}, num_noise=0, stepper=clode.Stepper.dormand_prince,
                                    device_id=0, platform_id=0)
    old = clode.get_log_level()
    try:
        clode.set_log_level(clode.LogLevel.trace)
        assert clode.get_log_level() == clode.LogLevel.trace
        sim.print_devices()
        captured = capfd.readouterr()
        assert "OpenCL" in captured.out and "gfx950" in captured.out
        assert captured.err == ""
        clode.set_log_level(clode.LogLevel.off)
        assert clode.get_log_level() == clode.LogLevel.off
        sim.print_devices()
        captured = capfd.readouterr()
        assert captured.out == "" and captured.err == ""
    finally:
        clode.set_log_level(old)


def test_vdp_dormand_prince_long(gpu_available):
    """test/test_vdp_long.py:10-11: the Van der Pol period for mu up to 100."""
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "van_der_pol.cl"), variables={"x": 1.0, "y": 1.0},
                                 parameters={"mu": 1.0}, observer=clode.Observer.threshold_2,
                                 stepper=clode.Stepper.dormand_prince, t_span=(0.0, 1000.0), max_store=20000,
                                 max_steps=20000)
    mus = [-1, 0, 0.01, 0.1, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0] + list(range(5, 101))
    sim.set_ensemble(parameters={"mu": mus})
    sim.features()
    periods = sim.get_observer_results().get_var_max("period")
    for i, mu in enumerate(mus):
        assert np.isclose(periods[i], vdp_period(mu), rtol=0.01, atol=1), (mu, periods[i])
