"""Full-size launches (the benchmark's 256 x 256 lactotroph grid and a 2^20
Lorenz ensemble) checked through size-independent properties: members are
independent, so a deterministic sample of members from the full launch must
equal the oracle's result for those members bit for bit (FP64), and every
member must report finite features and positive step counts."""
import os

import numpy as np
import pytest

import clode_amd as clode
from oracle import OracleEnsemble

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "clode_amd", "models")


def test_lactotroph_256x256_sample_bitexact(gpu_available):
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph.cl"),
                                 variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                 parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0}, aux=["ica"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                 single_precision=False, dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, t_span=(0.0, 1000.0), observer_x_up_thresh=0.3,
                                 observer_x_down_thresh=0.2)
    gk, gbk = np.linspace(0, 10, 256), np.linspace(0, 2, 256)
    GK, GBK = np.meshgrid(gk, gbk, indexing="ij")
    sim.set_ensemble(parameters={"gk": GK, "gbk": GBK})
    sim.enable_step_counter(True)
    sim.transient()
    sim.set_tspan((0.0, 3000.0))
    out = sim.features()
    F = out.to_ndarray().reshape(65536, -1)
    assert np.all(np.isfinite(F))
    assert np.all(sim.get_step_counts() > 0)
    idx = np.arange(7, 65536, 1021)
    o = OracleEnsemble("lactotroph", 4, 4, 1, stepper="dopri5", observer="thresh2", precision="double")
    o.set_solver_params(0.1, 1.0, 1e-8, 1e-6, 10 ** 8, 1000000, 1)
    o.set_observer_params(x_up_threshold=0.3, x_down_threshold=0.2)
    pars = np.vstack([1.5 * np.ones(idx.size), 3.0 * np.ones(idx.size), GBK.ravel()[idx], GK.ravel()[idx]])
    o.set_problem_data(np.tile([[-60.0], [0.0], [0.0], [0.1]], (1, idx.size)), pars)
    o.set_tspan((0.0, 1000.0))
    o.transient()
    o.shift_x0()
    o.set_tspan((0.0, 3000.0))
    o.features()
    ref = o.get("F").reshape(-1, idx.size).T
    assert np.array_equal(F[idx], ref)


@pytest.mark.parametrize("stepper", ["dopri5", "bs23"])
def test_lorenz_2p20_fp32c(gpu_available, stepper):
    n = 1 << 20
    rng = np.random.default_rng(5)
    x0 = rng.normal(1.0, 0.1, size=(n, 3))
    rho = np.linspace(0.0, 250.0, n)
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lorenz.cl"), variables={"x": 1.0, "y": 1.0, "z": 1.0},
                                 parameters={"sigma": 10.0, "rho": 28.0, "beta": 8 / 3},
                                 observer=clode.Observer.local_max, stepper=clode.Stepper(stepper),
                                 single_precision="fp32c", dt=0.01, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, t_span=(0.0, 50.0))
    sim.set_ensemble(variables=x0, parameters={"rho": rho})
    out = sim.features()
    F = out.to_ndarray().reshape(n, -1)
    assert np.all(np.isfinite(F))
    names = out.get_feature_names()
    ev = F[:, names.index("event count")]
    # rho < 1: the origin is stable, x decays monotonically or with no maxima after
    # the transient; rho > 24.74: chaos with many maxima
    assert np.median(ev[rho > 100]) > 10
    # non-chaotic members (rho < 13: no chaotic transient, stable equilibria),
    # 412 of them: the fp32c error against the FP64 oracle at the same
    # tolerance is compared with the method's own truncation error, measured
    # as the FP64 difference between rtol 1e-6 and rtol 1e-8.
    # * "mean x" (a time integral): within 3x that error for every member.
    # * "max z", "min y" (extrema over the accepted step points): these move by
    #   up to |x''| dt^2 / 8 whenever the step grid shifts, and a last-bit
    #   difference in one error norm near the accept/reject boundary shifts
    #   it (FP64 itself moves "max z" by up to 1e-2 on some members when rtol
    #   changes by 1e-4 relative).  Within 3x the truncation error for at
    #   least 95 % of the members, within 0.05 absolute for all of them.
    # * all three: median relative error below 1e-4.
    idx = np.arange(100, 54000, 131)
    refs = {}
    for rtol in (1e-6, 1e-8):
        o = OracleEnsemble("lorenz", 3, 3, stepper=stepper, observer="localmax", precision="double")
        o.set_solver_params(0.01, 1.0, rtol * 1e-2, rtol, 10 ** 8, 1000000, 1)
        o.set_observer_params()
        o.set_problem_data(x0[idx].T, np.vstack([10.0 * np.ones(idx.size), rho[idx], np.full(idx.size, 8 / 3)]))
        o.set_tspan((0.0, 50.0))
        o.features()
        refs[rtol] = o.get("F").reshape(-1, idx.size).T
    for key in ("mean x", "max z", "min y"):
        j = names.index(key)
        a, b, fine = F[idx, j], refs[1e-6][:, j], refs[1e-8][:, j]
        err = np.abs(a - b)
        within = err <= 3 * np.abs(b - fine) + 1e-5 * (1 + np.abs(b))
        if key == "mean x":
            assert np.all(within), key
        else:
            assert np.mean(within) >= 0.95, (key, float(np.mean(within)))
            assert np.all(err <= 0.05), (key, float(err.max()))
        assert float(np.median(err / np.abs(b))) < 1e-4, key


@pytest.mark.parametrize("stepper", ["dopri5", "bs23"])
def test_lorenz_2p20_fp64_occupancy_variant_sample_bitexact(gpu_available, stepper):
    """2^20 members are 16 waves per SIMD, more than the FP64 localmax
    features kernel keeps resident, so the grid launch runs the variant
    compiled for one more wave per SIMD; it must give the oracle's results
    bit for bit on a sample of members, like the default kernel."""
    n = 1 << 20
    rng = np.random.default_rng(8)
    x0 = rng.normal(1.0, 0.1, size=(n, 3))
    rho = np.linspace(0.0, 250.0, n)
    st = clode.Stepper.dormand_prince if stepper == "dopri5" else clode.Stepper.bs23
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lorenz.cl"), variables={"x": 1.0, "y": 1.0, "z": 1.0},
                                 parameters={"sigma": 10.0, "rho": 28.0, "beta": 8 / 3},
                                 observer=clode.Observer.local_max, stepper=st,
                                 single_precision=False, dt=0.01, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, t_span=(0.0, 20.0))
    sim.set_ensemble(variables=x0, parameters={"rho": rho})
    out = sim.features()
    assert sim.last_launch_variant() >= 2, "the occupancy variant did not run"
    F = out.to_ndarray().reshape(n, -1)
    idx = np.arange(3, n, 2039)
    o = OracleEnsemble("lorenz", 3, 3, stepper=stepper, observer="localmax", precision="double")
    o.set_solver_params(0.01, 1.0, 1e-8, 1e-6, 10 ** 8, 1000000, 1)
    o.set_observer_params()
    o.set_problem_data(x0[idx].T, np.vstack([10.0 * np.ones(idx.size), rho[idx], np.full(idx.size, 8 / 3)]))
    o.set_tspan((0.0, 20.0))
    o.features()
    assert np.array_equal(F[idx], o.get("F").reshape(-1, idx.size).T)


def test_lactotroph_2p17_occupancy_variant_sample_bitexact(gpu_available):
    """131,072 lactotroph members are two waves per SIMD; the FP64 thresh2
    features kernel keeps one resident, so the two-wave variant runs, bit for
    bit with the oracle on a sample."""
    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph.cl"),
                                 variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                 parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0}, aux=["ica"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                 single_precision=False, dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, t_span=(0.0, 300.0), observer_x_up_thresh=0.3,
                                 observer_x_down_thresh=0.2)
    GK, GBK = np.meshgrid(np.linspace(0, 10, 512), np.linspace(0, 2, 256), indexing="ij")
    sim.set_ensemble(parameters={"gk": GK, "gbk": GBK})
    out = sim.features()
    assert sim.last_launch_variant() >= 2, "the occupancy variant did not run"
    F = out.to_ndarray().reshape(GK.size, -1)
    idx = np.arange(11, GK.size, 509)
    o = OracleEnsemble("lactotroph", 4, 4, 1, stepper="dopri5", observer="thresh2", precision="double")
    o.set_solver_params(0.1, 1.0, 1e-8, 1e-6, 10 ** 8, 1000000, 1)
    o.set_observer_params(x_up_threshold=0.3, x_down_threshold=0.2)
    pars = np.vstack([1.5 * np.ones(idx.size), 3.0 * np.ones(idx.size), GBK.ravel()[idx], GK.ravel()[idx]])
    o.set_problem_data(np.tile([[-60.0], [0.0], [0.0], [0.1]], (1, idx.size)), pars)
    o.set_tspan((0.0, 300.0))
    o.features()
    assert np.array_equal(F[idx], o.get("F").reshape(-1, idx.size).T)


@pytest.mark.parametrize("precision", ["double", "fp32c"])
def test_lactotroph_256x256_replay_equals_integration(gpu_available, precision):
    """The full 256 x 256 grid with the features replay on (the warm-up's
    record of every member's steps, about 2,700 per member here) gives the
    integrating path's features, final states, step sizes and times bit for
    bit, in FP64 and in fp32c."""
    res = {}
    for budget in (0, 24 << 30):
        sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph.cl"),
                                     variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                     parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0}, aux=["ica"],
                                     observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                     single_precision=precision, dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                     max_steps=10 ** 8, t_span=(0.0, 1000.0), observer_x_up_thresh=0.3,
                                     observer_x_down_thresh=0.2)
        gk, gbk = np.linspace(0, 10, 256), np.linspace(0, 2, 256)
        GK, GBK = np.meshgrid(gk, gbk, indexing="ij")
        sim.set_ensemble(parameters={"gk": GK, "gbk": GBK})
        sim.set_replay_budget(budget)
        sim.transient()
        sim.set_tspan((0.0, 3000.0))
        F = sim.features(update_x0=False).to_ndarray()
        res[budget] = (F, sim.get_final_state(), sim.get_dt(), sim.replay_info())
        sim.set_replay_budget(0)
    (f0, x0, dt0, info0), (f1, x1, dt1, info1) = res[0], res[24 << 30]
    assert info0 == (0, False) and info1[1] and info1[0] >= 3000
    assert np.array_equal(f0, f1, equal_nan=True)
    assert np.array_equal(x0, x1) and np.array_equal(dt0, dt1)
