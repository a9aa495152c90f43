"""GPU parity on randomly drawn configurations, bit for bit against the CPU
oracle: a fixed list of seeds draws the model, stepper, observer, precision,
ensemble size, solver settings and time span of each case.  Besides widening
the fixed parity matrix, the draws put the time spans where the controller's
minimum step takes its other branches (negative times, spans that cross
zero, times just below a power of two, large offsets), with per-member spans
on some cases, and the observers' event lists on.  Every seed in the second
half draws a noise model (Euler-Maruyama, seeded stream, drawn seed) with any
of the six observers, and the noise stream's state after the run must agree
as well.
"""
import numpy as np
import pytest

from helpers import Pair

pytestmark = pytest.mark.gpu

MODELS = {
    "van_der_pol": (2, 1, lambda rng, n: (np.vstack([rng.uniform(1.0, 2.5, n), rng.uniform(-1, 1, n)]),
                                         rng.uniform(0.1, 4.0, n)[None, :]), 40.0),
    "lorenz": (3, 3, lambda rng, n: (rng.normal(1.0, 0.1, (3, n)),
                                     np.vstack([10.0 * np.ones(n), rng.uniform(0, 60, n), np.full(n, 8 / 3)])), 10.0),
    "lactotroph": (4, 4, lambda rng, n: (np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n)),
                                         np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), rng.uniform(0, 2, n),
                                                    rng.uniform(0, 10, n)])), 600.0),
    # noise models (one Wiener increment per step)
    "lactotroph_noise": (4, 4, lambda rng, n: (np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n)),
                                               np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), rng.uniform(0, 2, n),
                                                          rng.uniform(0, 4, n)])), 400.0),
    "ornstein_uhlenbeck": (1, 3, lambda rng, n: (rng.normal(0.0, 1.0, (1, n)),
                                                 np.vstack([rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n),
                                                            rng.uniform(0.1, 1, n)])), 50.0),
}
OBSERVERS = ["thresh2", "localmax", "nhood2", "basicall", "basic", "nhood1"]
NOISY = ("lactotroph_noise", "ornstein_uhlenbeck")
ADAPTIVE = ("bs23", "dopri5", "he12")
FIXED = ("euler", "heun", "rk4")


def draw(seed):
    rng = np.random.default_rng(seed)
    if seed < 24:
        model = ["van_der_pol", "lorenz", "lactotroph"][seed % 3]
        stepper = (ADAPTIVE + FIXED)[rng.integers(6)]
        if stepper in FIXED and model == "lactotroph":
            stepper = "dopri5"  # stiff spikes: a fixed step of this size diverges
        observer = OBSERVERS[rng.integers(4)]
    else:
        model = ["van_der_pol", "lorenz", "lactotroph", NOISY[0], NOISY[1], NOISY[seed % 2]][seed % 6]
        stepper = "seuler" if model in NOISY else (ADAPTIVE + FIXED)[rng.integers(6)]
        if stepper in FIXED and model == "lactotroph":
            stepper = "dopri5"
        observer = OBSERVERS[rng.integers(6)]
    prec = ["double", "single"][rng.integers(2)]
    n = int(rng.integers(33, 150))
    _, _, data, length = MODELS[model]
    start = [0.0, -0.5 * length, 1.0 - 2.0 ** -20, 1000.0, -3.0, 0.25][rng.integers(6)]
    # a tolerance the precision can meet (float carries about 7 digits)
    rtol = float(10 ** (rng.uniform(-8, -4) if prec == "double" else rng.uniform(-5, -3)))
    dt = 0.05 if stepper == "seuler" else 0.01 if stepper in FIXED else float(10 ** rng.uniform(-3, -1))
    # Before t = 0 the reference's minimum step 16 |fabs(nextafter(t, 1.1 t1)) - t|
    # is about 2 |t|, above dtmax: every adaptive step then fails at hmin
    # without advancing, until max_steps ends the member (reproduced on both
    # sides); a smaller step budget keeps those cases short.
    max_steps = 20000 if (start < 0 and stepper in ADAPTIVE) else 200000
    sp = (dt, 1.0, rtol * 1e-2, rtol, max_steps, 10 ** 6, 1)
    per_member = bool(rng.integers(2))
    return model, stepper, observer, prec, n, data(rng, n), (start, start + length), sp, per_member, rng


@pytest.mark.parametrize("seed", list(range(48)))
def test_random_configuration_bitexact(gpu_available, seed):
    model, stepper, observer, prec, n, (x0, pars), tspan, sp, per_member, rng = draw(seed)
    print(f"seed {seed}: {model} {stepper} {observer} {prec} n={n} tspan={tspan} sp={sp} "
          f"per_member={per_member}", flush=True)
    p = Pair(model, stepper, observer, prec, obs=dict(max_event_timestamps=int(rng.integers(0, 4))), sp=sp)
    p.set_problem_data(x0, pars)
    if model in NOISY:
        p.seed(int(rng.integers(0, 2 ** 31 - 1)))
    p.set_tspan(*tspan)
    if per_member:
        t0 = tspan[0] + rng.uniform(-1.0, 1.0, n)
        p.set_member_tspan(np.vstack([t0, t0 + (tspan[1] - tspan[0]) * rng.uniform(0.2, 1.0, n)]))
    p.transient()
    p.assert_bitexact(("xf", "dt", "tf"))
    p.shift_x0()
    p.features()
    p.assert_bitexact(("F", "xf", "dt", "tf"))
    if model in NOISY:
        assert np.array_equal(np.array(p.gpu.get_rng_state(), dtype=np.uint64), p.orc.get_rng())


def draw_trajectory(seed):
    rng = np.random.default_rng(1000 + seed)
    model = ["van_der_pol", "lorenz", "lactotroph", "lactotroph_noise", "ornstein_uhlenbeck"][seed % 5]
    if model in NOISY:
        stepper = "seuler"
    else:
        stepper = (ADAPTIVE + FIXED)[rng.integers(6)]
        if stepper in FIXED and model == "lactotroph":
            stepper = "rk4"  # with the small fixed step below
    prec = ["double", "single"][rng.integers(2)]
    n = int(rng.integers(17, 140))
    _, _, data, length = MODELS[model]
    start = [0.0, 1000.0, 0.25, 1.0 - 2.0 ** -20][rng.integers(4)]
    rtol = float(10 ** (rng.uniform(-8, -4) if prec == "double" else rng.uniform(-5, -3)))
    dt = {"seuler": 0.05, "rk4": 0.01 if model != "lactotroph" else 0.002}.get(stepper, 0.01)
    if stepper in ADAPTIVE:
        dt = float(10 ** rng.uniform(-3, -1))
    nout = int(rng.integers(1, 6))
    max_store = int(rng.integers(1, 300))
    sp = (dt, 1.0, rtol * 1e-2, rtol, 200000, max_store, nout)
    return model, stepper, prec, n, data(rng, n), (start, start + length * rng.uniform(0.1, 0.5)), sp, rng


@pytest.mark.parametrize("seed", list(range(16)))
def test_random_trajectory_bitexact(gpu_available, seed):
    """Randomly drawn trajectory launches (model, stepper, precision, output
    stride nout, max_store, time origin, noise with a drawn seed): every
    stored sample (t, x, dx, aux), the stored counts and the final state agree
    with the oracle bit for bit."""
    model, stepper, prec, n, (x0, pars), tspan, sp, rng = draw_trajectory(seed)
    print(f"trajectory seed {seed}: {model} {stepper} {prec} n={n} tspan={tspan} sp={sp}", flush=True)
    p = Pair(model, stepper, "", prec, kind="trajectory", sp=sp)
    p.set_problem_data(x0, pars)
    if model in NOISY:
        p.seed(int(rng.integers(0, 2 ** 31 - 1)))
    p.set_tspan(*tspan)
    p.trajectory()
    p.assert_trajectory_bitexact()
    if model in NOISY:
        assert np.array_equal(np.array(p.gpu.get_rng_state(), dtype=np.uint64), p.orc.get_rng())
