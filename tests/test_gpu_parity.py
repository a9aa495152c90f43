"""GPU parity: the HIP kernels (through the C ABI via the pybind11 module)
against the CPU oracle on the same inputs, bit for bit.

Plain single and double precision run the reference's arithmetic with the
same fused multiply-adds and the same deterministic exp/log/pow on both
sides, so every state, step size, time, feature and RNG word must be equal,
not merely close.  Ensemble sizes are deliberately not multiples of the
64-lane wavefront.
"""
import math
import os

import numpy as np
import pytest

from helpers import MODEL_INFO, MODELS, OBS_DEFAULTS, Pair, wrapper

pytestmark = pytest.mark.gpu

FIXED = (0.01, 1.0, 1e-6, 1e-3, 1000000, 1000000, 1)
ADAPT = (0.1, 1.0, 1e-8, 1e-6, 10000000, 1000000, 1)


def vdp_data(n):
    mu = np.linspace(0.1, 5.0, n)
    return np.vstack([2.0 * np.ones(n), np.zeros(n)]), mu[None, :]


def lac_data(n, noise=False):
    g = np.linspace(0.0, 10.0, n)
    b = np.linspace(2.0, 0.0, n)
    x0 = np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n))
    last = np.full(n, 2.0) if noise else g  # sigma or gk
    pars = np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), b, last])
    return x0, pars


def lorenz_data(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.normal(1.0, 0.1, (3, n)), np.vstack([10.0 * np.ones(n), np.linspace(0, 250, n), np.full(n, 8 / 3)])


CASES = []
for prec in ("double", "single"):
    for stepper in ("euler", "heun", "rk4"):
        for observer in ("thresh2", "localmax", "nhood2"):
            CASES.append(("van_der_pol", stepper, observer, prec, 97, (0.0, 60.0), FIXED))
    for stepper in ("bs23", "dopri5"):
        CASES.append(("lactotroph", stepper, "thresh2", prec, 70, (0.0, 1500.0), ADAPT))
        CASES.append(("lorenz", stepper, "localmax", prec, 130, (0.0, 40.0), ADAPT))
        CASES.append(("lactotroph", stepper, "nhood2", prec, 33, (0.0, 1000.0), ADAPT))
    for observer in ("thresh2", "localmax", "nhood2"):
        CASES.append(("lactotroph_noise", "seuler", observer, prec, 65, (0.0, 400.0),
                      (0.05, 1.0, 1e-6, 1e-3, 1000000, 1000000, 1)))
    CASES.append(("van_der_pol", "rk4", "basicall", prec, 97, (0.0, 60.0), FIXED))
    CASES.append(("lorenz", "bs23", "basic", prec, 130, (0.0, 40.0), ADAPT))
    CASES.append(("lactotroph", "dopri5", "nhood1", prec, 70, (0.0, 1500.0), ADAPT))
    # Heun-Euler 1(2) (adaptive_he12.clh), a low-order method: looser tolerance
    CASES.append(("van_der_pol", "he12", "thresh2", prec, 97, (0.0, 60.0), (0.01, 1.0, 1e-6, 1e-4, 10000000, 1000000, 1)))
    CASES.append(("lorenz", "he12", "localmax", prec, 130, (0.0, 20.0), (0.01, 1.0, 1e-6, 1e-4, 10000000, 1000000, 1)))
    CASES.append(("lactotroph_noise", "seuler", "basicall", prec, 65, (0.0, 400.0),
                  (0.05, 1.0, 1e-6, 1e-3, 1000000, 1000000, 1)))


def data_for(model, n):
    if model == "van_der_pol":
        return vdp_data(n)
    if model == "lactotroph":
        return lac_data(n)
    if model == "lactotroph_noise":
        return lac_data(n, noise=True)
    return lorenz_data(n)


@pytest.mark.parametrize("model,stepper,observer,prec,n,tspan,sp", CASES,
                         ids=[f"{c[0]}-{c[1]}-{c[2]}-{c[3]}" for c in CASES])
def test_transient_then_features_bitexact(gpu_available, model, stepper, observer, prec, n, tspan, sp):
    p = Pair(model, stepper, observer, prec, obs=dict(max_event_timestamps=3), sp=sp)
    x0, pars = data_for(model, n)
    p.set_problem_data(x0, pars)
    p.seed(1234)
    p.set_tspan(*tspan)
    p.transient()
    p.assert_bitexact(("xf", "dt", "tf"))
    p.shift_x0()
    p.features()
    res = p.assert_bitexact(("F", "xf", "dt", "tf"))
    assert res["F"][1] == n * p.gpu.get_n_features()
    # the oracle and the device consumed the same noise
    assert np.array_equal(np.array(p.gpu.get_xf()), p.orc.get("xf"))


@pytest.mark.parametrize("prec", ["double", "single"])
def test_observer_continuation_and_reinit(gpu_available, prec):
    """features() twice continues the observer (time-shifted data, counters
    keep growing); features(True) re-initialises (CLODEfeatures.cpp:215-229)."""
    p = Pair("van_der_pol", "rk4", "thresh2", prec, obs=dict(max_event_timestamps=4), sp=FIXED)
    x0, pars = vdp_data(70)
    p.set_problem_data(x0, pars)
    p.set_tspan(0.0, 40.0)
    p.features()
    p.shift_x0()
    p.features()  # continued
    p.assert_bitexact(("F", "xf"))
    F1 = np.array(p.gpu.get_f()).reshape(-1, 70)
    assert p.gpu.is_observer_initialized()
    p.shift_x0()
    p.features(True)  # re-initialised
    p.assert_bitexact(("F", "xf"))
    F2 = np.array(p.gpu.get_f()).reshape(-1, 70)
    names = p.gpu.get_feature_names()
    sc = names.index("step count")
    assert np.all(F1[sc] > F2[sc])  # continuation accumulates steps, reinit restarts


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("stepper", ["rk4", "dopri5", "seuler"])
def test_trajectory_bitexact(gpu_available, prec, stepper):
    model = "lactotroph_noise" if stepper == "seuler" else "lactotroph"
    sp = (0.05, 1.0, 1e-8, 1e-6, 10000000, 400, 3)  # nout 3, max_store 400
    p = Pair(model, stepper, "", prec, kind="trajectory", sp=sp)
    x0, pars = lac_data(77, noise=(model == "lactotroph_noise"))
    p.set_problem_data(x0, pars)
    p.seed(99)
    p.set_tspan(0.0, 300.0)
    p.trajectory()
    assert p.assert_trajectory_bitexact() > 77
    ns = np.array(p.gpu.get_n_stored())
    assert ns.max() <= 401


def test_trajectory_max_store_truncation(gpu_available):
    """The store loop stops once max_store samples exist; nStored counts the
    last step even though sample max_store is never written (reference H9)."""
    sp = (0.01, 1.0, 1e-6, 1e-3, 10000000, 50, 1)
    p = Pair("van_der_pol", "rk4", "", "double", kind="trajectory", sp=sp)
    x0, pars = vdp_data(5)
    p.set_problem_data(x0, pars)
    p.set_tspan(0.0, 100.0)
    p.trajectory()
    p.assert_trajectory_bitexact()
    assert np.all(np.array(p.gpu.get_n_stored()) == 50)
    tf = np.array(p.gpu.get_tf())
    assert np.allclose(tf, 0.5, atol=1e-9)  # 50 steps of 0.01


@pytest.mark.parametrize("max_store", [0, 1, 2])
def test_trajectory_tiny_max_store(gpu_available, max_store):
    """max_store 0 writes no sample and takes no step (the loop runs while
    storeix < max_store); 1 and 2 stop after one and two steps."""
    sp = (0.01, 1.0, 1e-6, 1e-3, 10000000, max_store, 1)
    p = Pair("van_der_pol", "rk4", "", "double", kind="trajectory", sp=sp)
    x0, pars = vdp_data(70)
    p.set_problem_data(x0, pars)
    p.set_tspan(0.0, 10.0)
    p.trajectory()
    p.assert_bitexact(("xf", "tf"))
    ns = np.array(p.gpu.get_n_stored())
    assert np.all(ns == max_store)
    if max_store == 0:
        assert np.array_equal(np.array(p.gpu.get_xf()), x0.ravel())
    else:
        p.assert_trajectory_bitexact()


def test_seeded_noise_stream(gpu_available):
    p = Pair("ornstein_uhlenbeck", "seuler", "", "double", kind="base",
             sp=(0.01, 0.01, 1e-6, 1e-3, 1000000, 1000000, 1))
    n = 300
    p.set_problem_data(np.zeros((1, n)), np.vstack([np.ones(n), np.ones(n), 0.5 * np.ones(n)]))
    p.set_tspan(0.0, 5.0)
    p.seed(2024)
    p.transient()
    p.assert_bitexact(("xf",))
    # identical seed -> identical result; the RNG words after the run agree too
    xf1 = np.array(p.gpu.get_xf())
    p.seed(2024)
    p.gpu.transient()
    assert np.array_equal(np.array(p.gpu.get_xf()), xf1)
    p.seed(2025)
    p.gpu.transient()
    assert not np.array_equal(np.array(p.gpu.get_xf()), xf1)
    # OU stationary law: x ~ N(mu, sigma^2/(2 theta)) after t >> 1/theta
    p.set_tspan(0.0, 20.0)
    p.seed(7)
    p.gpu.transient()
    x = np.array(p.gpu.get_xf())
    assert abs(x.mean() - 1.0) < 0.08 and abs(x.var() - 0.125) < 0.04


def test_dt_refill_semantics(gpu_available):
    """The device step buffer is refilled with sp.dt only when nPts changes
    (CLODE.cpp:184-202, 377-403); set_dt_refresh(True) refills on every
    set_solver_params."""
    p = Pair("decay", "euler", "", "double", kind="base", sp=(0.1, 1.0, 1e-6, 1e-3, 1000, 1000, 1))
    p.set_problem_data(np.ones((1, 3)), np.ones((1, 3)))
    p.set_solver_params((0.05, 1.0, 1e-6, 1e-3, 1000, 1000, 1))
    assert np.all(np.array(p.gpu.get_dt()) == 0.1)
    p.gpu.set_dt_refresh(True)
    p.set_solver_params((0.05, 1.0, 1e-6, 1e-3, 1000, 1000, 1))
    assert np.all(np.array(p.gpu.get_dt()) == 0.05)
    p.set_problem_data(np.ones((1, 4)), np.ones((1, 4)))  # nPts change refills
    assert np.all(np.array(p.gpu.get_dt()) == 0.05)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 257])
def test_ensemble_sizes_and_max_steps(gpu_available, n):
    sp = (0.01, 1.0, 1e-6, 1e-3, 777, 1000000, 1)  # max_steps truncates
    p = Pair("van_der_pol", "rk4", "thresh2", "double", sp=sp)
    x0, pars = vdp_data(n)
    p.set_problem_data(x0, pars)
    p.set_tspan(0.0, 100.0)
    p.features()
    p.assert_bitexact(("F", "xf", "tf"))
    names = p.gpu.get_feature_names()
    F = np.array(p.gpu.get_f()).reshape(-1, n)
    assert np.all(F[names.index("step count")] == 777)


def test_terminal_event_count(gpu_available):
    p = Pair("harmonic", "rk4", "localmax", "double", obs=dict(max_event_count=3, max_event_timestamps=5), sp=FIXED)
    n = 40
    p.set_problem_data(np.vstack([np.ones(n), np.zeros(n)]), np.linspace(1.0, 3.0, n)[None, :])
    p.set_tspan(0.0, 100.0)
    p.features()
    p.assert_bitexact(("F", "xf", "tf"))
    names = p.gpu.get_feature_names()
    F = np.array(p.gpu.get_f()).reshape(-1, n)
    assert np.all(F[names.index("event count")] == 3)
    assert np.all(np.array(p.gpu.get_tf()) < 100.0)  # stopped at the 3rd maximum
    omega = np.linspace(1.0, 3.0, n)
    # x = cos(omega t) starts at a maximum; detected maxima at 2 pi k / omega, k = 1, 2, 3
    assert np.allclose(F[names.index("localmax event time 1")], 4 * math.pi / omega, atol=0.02)


def test_step_counter_matches_oracle(gpu_available):
    p = Pair("lactotroph", "dopri5", "thresh2", "double", sp=ADAPT)
    x0, pars = lac_data(50)
    p.set_problem_data(x0, pars)
    p.gpu.set_step_counter(True)
    p.set_tspan(0.0, 800.0)
    p.transient()
    tr = np.array(p.gpu.get_step_counts(), dtype=np.int64).sum()
    assert tr == p.orc.last_steps
    p.gpu.reset_step_counts()
    p.features()
    total = np.array(p.gpu.get_step_counts(), dtype=np.int64).sum()
    assert total == p.orc.last_steps + p.orc.last_warm_steps


@pytest.mark.parametrize("prec", ["double", "fp32c"])
def test_large_ensemble_variant_bitexact(gpu_available, prec):
    """Ensembles of at least two waves per SIMD run the kernels compiled for
    two waves per SIMD when the default ones fit only one (lactotroph
    features in FP64 and fp32c): same arithmetic, so the features of 262,144
    members equal those of the same members run as four ensembles of 65,536
    (one wave per SIMD: the default kernels)."""
    w = wrapper()
    vars_, pars, aux, noise = MODEL_INFO["lactotroph"]
    pi = w.ProblemInfo(os.path.join(MODELS, "lactotroph.cl"), vars_, pars, aux, noise)
    op = w.ObserverParams(**OBS_DEFAULTS)

    def run(x0, p):
        s = w.FeatureSimulatorBase(pi, "dopri5", "thresh2", op, prec, w.OpenCLResource(), "")
        s.build_cl()
        s.set_solver_params(w.SolverParams(0.1, 1.0, 1e-8, 1e-6, 1000000, 1000000, 1))
        s.set_problem_data(np.ascontiguousarray(x0).ravel(), np.ascontiguousarray(p).ravel())
        s.set_tspan([0.0, 300.0])
        s.features(True)
        return np.array(s.get_f()).reshape(-1, x0.shape[1]), np.array(s.get_xf())

    n = 1 << 18
    x0, p = lac_data(n)
    F, xf = run(x0, p)
    q = n // 4
    parts = [run(x0[:, k * q:(k + 1) * q], p[:, k * q:(k + 1) * q]) for k in range(4)]
    assert np.array_equal(F, np.concatenate([Fk for Fk, _ in parts], axis=1))
    assert np.array_equal(xf.reshape(4, n), np.concatenate([xk.reshape(4, q) for _, xk in parts], axis=1))


@pytest.mark.parametrize("stepper", ["bs23", "dopri5"])
def test_controller_long_retry_and_stall(stepper):
    """The adaptive controller's two ends of the retry loop, on the device
    against the oracle (FP64) bit for bit: x' = -lambda x with dt = 1 from
    t = 0 and lambda ~ 1e31 needs more than 100 halvings before the first
    step is accepted (members 0-2); from t0 = 1 - 13 2^-53 with lambda = 1e40
    every trial fails and the trial at hmin rounds to a longer step, so it
    would repeat forever (members 3-5): each step fails at hmin and leaves t
    and x unchanged.  (Float modes overflow on these right-hand sides.)"""
    sp = (1.0, 1.0, 1e-12, 1e-6, 3, 1000, 1)
    lam = np.array([1e31, 2e31, 5e30])
    for t0, lams in ((0.0, lam), (1.0 - 13 * 2.0 ** -53, np.full(3, 1e40))):
        print(f"controller case {stepper} t0={t0!r} lambda={lams}", flush=True)
        p = Pair("decay", stepper, "", "double", kind="base", sp=sp)
        p.set_problem_data(np.ones((1, 3)), lams[None, :])
        p.set_tspan(t0, 2.0)
        p.transient()
        p.assert_bitexact(("xf", "dt", "tf"))
        tf = p.gpu_get("tf")
        if t0 == 0.0:
            assert np.all(tf > 0) and np.all(tf < 1e-29)
        else:
            assert np.all(tf == t0) and np.all(p.gpu_get("dt") == 2.0 ** -49)


@pytest.mark.parametrize("observer", ["thresh2", "localmax", "nhood2"])
def test_continuation_equals_one_long_run(observer):
    """Observer continuation on the device, with the exact setup that pins it
    in the oracle (tests/test_oracle_continuation.py): harmonic oscillator,
    rk4, dt = 1/64, spans ending half a step past a multiple of dt.  The
    continued features (initialise over the whole span, features over
    (0, 17), shift_x0, features over (0, 23)) equal one features() over
    (0, 40) bit for bit on the device, and both equal the oracle."""
    dt = 1.0 / 64
    sp = (dt, 1.0, 1e-6, 1e-3, 10 ** 7, 1000, 1)
    obs = dict(max_event_count=12)
    n = 70
    x0 = np.vstack([np.ones(n), np.zeros(n)])
    om = np.linspace(1.5, 3.1, n)[None, :]
    whole = Pair("harmonic", "rk4", observer, "double", sp=sp, obs=obs)
    whole.set_problem_data(x0, om)
    whole.set_tspan(0.0, 40.0 - dt / 2)
    whole.features(True)
    whole.assert_bitexact(("F", "xf", "tf"))
    split = Pair("harmonic", "rk4", observer, "double", sp=sp, obs=obs)
    split.set_problem_data(x0, om)
    split.set_tspan(0.0, 40.0 - dt / 2)
    split.initialize_observer()
    split.set_tspan(0.0, 17.0 - dt / 2)
    split.features()
    split.shift_x0()
    split.set_tspan(0.0, 23.0 - dt / 2)
    split.features()
    split.assert_bitexact(("F", "xf", "tf"))
    assert np.array_equal(split.gpu_get("F"), whole.gpu_get("F"))
    assert np.array_equal(split.gpu_get("tf") + 17.0, whole.gpu_get("tf"))


@pytest.mark.parametrize("model,stepper,prec", [("lactotroph", "dopri5", "double"), ("lactotroph", "dopri5", "fp32c"),
                                                ("lactotroph_noise", "seuler", "double"),
                                                ("lactotroph_noise", "seuler", "fp32c")])
def test_transient_features_one_launch(model, stepper, prec):
    """clode_transient_features (transient, x0 <- xf, warm-up and features of
    each member back to back in one launch) against the separate calls
    transient(), shift_x0(), features(true) with the same spans: features,
    states, step sizes, end times and noise words equal bit for bit; in FP64
    both also equal the oracle."""
    n = 130
    sp = ADAPT if stepper == "dopri5" else (0.05, 1.0, 1e-6, 1e-3, 10 ** 7, 10, 1)
    x0, pars = lac_data(n, noise=(model == "lactotroph_noise"))
    out = []
    for fused in (False, True):
        print(f"transient_features {model} {prec} fused={fused}", flush=True)
        p = Pair(model, stepper, "thresh2", prec, sp=sp, oracle=(prec == "double"))
        p.set_problem_data(x0, pars)
        p.seed(31)
        if fused:
            p.gpu.transient_features([0.0, 700.0], [0.0, 1500.0])
            if p.orc:
                p.orc.set_tspan((0.0, 700.0))
                p.orc.transient()
                p.orc.shift_x0()
                p.orc.set_tspan((0.0, 1500.0))
                p.orc.features(1)
        else:
            p.set_tspan(0.0, 700.0)
            p.transient()
            p.shift_x0()
            p.set_tspan(0.0, 1500.0)
            p.features(True)
        if p.orc:
            p.assert_bitexact(("F", "xf", "dt", "tf", "x0"))
        assert tuple(p.gpu.get_tspan()) == (0.0, 1500.0)
        out.append([p.gpu_get(k) for k in ("F", "xf", "x0", "dt", "tf")] + [np.array(p.gpu.get_rng_state())])
    for a, b in zip(*out):
        assert np.array_equal(a, b)
