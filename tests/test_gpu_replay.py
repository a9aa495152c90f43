"""Features replay: the two-pass observers' warm-up pass records every
member's accepted steps in device memory and the features pass reads them
back instead of integrating them again.  The results must be those of the
integrating path bit for bit, so every case here runs the replaying GPU
simulator against the CPU oracle (which integrates both passes, as the
reference's CLODEfeatures::initializeObserver and features do,
clode/cpp/CLODEfeatures.cpp:180-259): the record's capacity limit (members
that integrate on from their last record), a maximum event count that ends
members in the middle of their record, per-member spans, the explicit
initialize_observer() + features(False) sequence, a record invalidated
by a change between the two calls, and the randomly drawn parity
configurations that have a two-pass observer.
"""
import numpy as np
import pytest

from helpers import Pair

pytestmark = pytest.mark.gpu

LACTO = (lambda rng, n: (np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n)),
                         np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), rng.uniform(0, 2, n),
                                    rng.uniform(0, 10, n)])))
DATA = {
    "lactotroph": (LACTO, (0.0, 600.0), 11),
    "van_der_pol": (lambda rng, n: (np.vstack([rng.uniform(1.0, 2.5, n), rng.uniform(-1, 1, n)]),
                                    rng.uniform(0.1, 4.0, n)[None, :]), (0.0, 60.0), 2 + 4),
    "lorenz": (lambda rng, n: (rng.normal(1.0, 0.1, (3, n)),
                               np.vstack([10.0 * np.ones(n), rng.uniform(0, 60, n), np.full(n, 8 / 3)])),
               (0.0, 12.0), 2 + 6),
}


def record_doubles(model, precision):
    nv = {"lactotroph": 4, "van_der_pol": 2, "lorenz": 3}[model]
    return DATA[model][2] + (nv if precision == "fp32c" else 0)


def make_pair(model, stepper, observer, precision, n, seed, obs=None, dt=0.05, oracle=True):
    sp = (dt, 1.0, 1e-8, 1e-6, 1000000, 1000000, 1) if precision == "double" else \
        (dt, 1.0, 1e-5, 1e-4, 1000000, 1000000, 1)
    p = Pair(model, stepper, observer, precision, obs=obs, sp=sp, oracle=oracle)
    rng = np.random.default_rng(seed)
    x0, pars = DATA[model][0](rng, n)
    p.set_problem_data(x0, pars)
    p.set_tspan(*DATA[model][1])
    return p, rng


class Twin:
    """fp32c has no oracle (it is not the reference's arithmetic): the
    replaying simulator is checked against a second GPU simulator that
    integrates both passes, on the same calls."""

    def __init__(self, *args, **kw):
        self.p, _ = make_pair(*args, oracle=False, **kw)
        self.q, _ = make_pair(*args, oracle=False, **kw)
        self.gpu = self.p.gpu

    def __getattr__(self, name):
        if name in ("transient", "shift_x0", "features", "initialize_observer"):
            def both(*a):
                getattr(self.p, name)(*a)
                getattr(self.q, name)(*a)
            return both
        raise AttributeError(name)

    def assert_bitexact(self, names):
        for name in names:
            a, b = self.p.gpu_get(name), self.q.gpu_get(name)
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            assert same.all(), f"{name}: {int(same.sum())}/{a.size} bit-exact"


CASES = [
    ("lactotroph", "dopri5", "thresh2", "double"),
    ("lactotroph", "dopri5", "thresh2", "single"),
    ("lactotroph", "dopri5", "thresh2", "fp32c"),
    ("lactotroph", "bs23", "thresh2", "double"),
    ("van_der_pol", "rk4", "thresh2", "double"),
    ("van_der_pol", "heun", "nhood2", "fp32c"),
    ("lorenz", "bs23", "nhood2", "double"),
    ("lorenz", "dopri5", "nhood2", "single"),
]


@pytest.mark.parametrize("model,stepper,observer,precision", CASES)
@pytest.mark.parametrize("capacity", ["all", "16"])
def test_replay_bitexact(gpu_available, model, stepper, observer, precision, capacity):
    """transient, shift_x0, features(True) with the replay on: the features
    pass replays the warm-up record (all steps, or only the first 16 per
    member, after which every member integrates on), and F, xf, dt, tf equal
    the oracle's (in fp32c: those of the integrating GPU path); the continued
    features() that follows has no record and integrates."""
    n = 97
    print(f"replay {model} {stepper} {observer} {precision} capacity={capacity}", flush=True)
    p = Twin(model, stepper, observer, precision, n, seed=5) if precision == "fp32c" else \
        make_pair(model, stepper, observer, precision, n, seed=5)[0]
    rec = record_doubles(model, precision) * 8
    p.gpu.set_replay_budget((1 << 30) if capacity == "all" else 16 * n * rec + rec)
    p.transient()
    p.assert_bitexact(("xf", "dt", "tf"))
    p.shift_x0()
    p.features(True)
    cap, replayed = p.gpu.get_replay_info()
    assert replayed and (cap == 16 if capacity == "16" else cap > 1000)
    p.assert_bitexact(("F", "xf", "dt", "tf"))
    p.features()  # continuation: observer data carried over, no record
    assert not p.gpu.get_replay_info()[1]
    p.assert_bitexact(("F", "xf", "dt", "tf"))


def test_replay_terminal_inside_record(gpu_available):
    """max_event_count 3 ends most lactotroph members at their third event,
    long before the end of their warm-up record: the stored final state, step
    size and time are those of the replayed step."""
    p, _ = make_pair("lactotroph", "dopri5", "thresh2", "double", 120, seed=8, obs=dict(max_event_count=3))
    p.gpu.set_replay_budget(1 << 30)
    p.features(True)
    assert p.gpu.get_replay_info()[1]
    p.assert_bitexact(("F", "xf", "dt", "tf"))
    tf = p.gpu_get("tf")
    assert np.mean(tf < 0.9 * DATA["lactotroph"][1][1]) > 0.5  # most members stopped early


def test_replay_member_spans_and_explicit_initialize(gpu_available):
    """Per-member spans, and the warm-up run by an explicit
    initialize_observer() followed by features(False)."""
    n = 130
    p, rng = make_pair("lactotroph", "dopri5", "thresh2", "double", n, seed=9)
    p.gpu.set_replay_budget(1 << 30)
    t0 = rng.uniform(-1.0, 1.0, n)
    p.set_member_tspan(np.vstack([t0, t0 + 600.0 * rng.uniform(0.2, 1.0, n)]))
    p.initialize_observer()
    p.features(False)
    assert p.gpu.get_replay_info()[1]
    p.assert_bitexact(("F", "xf", "dt", "tf"))


def test_replay_invalidated_by_a_change(gpu_available):
    """A change of the parameters between initialize_observer() and
    features(False) leaves the record unused: the features pass integrates,
    with the new parameters and the warm-up's observer data, as the oracle
    does."""
    n = 80
    p, rng = make_pair("lactotroph", "dopri5", "thresh2", "double", n, seed=10)
    p.gpu.set_replay_budget(1 << 30)
    p.initialize_observer()
    x0, pars = LACTO(rng, n)
    p.set_problem_data(x0, pars)
    p.features(False)
    assert not p.gpu.get_replay_info()[1]
    p.assert_bitexact(("F", "xf", "dt", "tf"))


def test_replay_step_counts(gpu_available):
    """Replayed steps are not counted as member-steps: with the replay the
    features pass counts only the steps it integrated past the record (at
    most a few per member), without it every step."""
    n = 64
    counts = {}
    for budget in (0, 1 << 30):
        p, _ = make_pair("lactotroph", "dopri5", "thresh2", "double", n, seed=11, obs=dict(max_event_count=100000))
        p.orc = None
        p.gpu.set_replay_budget(budget)
        p.gpu.set_step_counter(True)
        p.gpu.initialize_observer()
        warm = np.array(p.gpu.get_step_counts(), dtype=np.int64)
        p.gpu.features(False)
        counts[budget] = (warm, np.array(p.gpu.get_step_counts(), dtype=np.int64) - warm,
                          np.array(p.gpu.get_f()))
    warm0, feat0, f0 = counts[0]
    warm1, feat1, f1 = counts[1 << 30]
    assert np.array_equal(warm0, warm1) and np.array_equal(f0, f1)
    assert np.all(feat0 >= warm0) and np.all(feat1 == feat0 - warm0)


def test_replay_not_used_with_noise(gpu_available):
    """A system with noise integrates both passes (the record holds no noise
    stream position): no replay, results and stream state as the oracle's."""
    n = 70
    p = Pair("lactotroph_noise", "seuler", "thresh2", "double", sp=(0.05, 1.0, 1e-8, 1e-6, 1000000, 1000000, 1))
    rng = np.random.default_rng(12)
    p.set_problem_data(np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n)),
                       np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), rng.uniform(0, 2, n), rng.uniform(0, 4, n)]))
    p.seed(1234)
    p.set_tspan(0.0, 300.0)
    p.gpu.set_replay_budget(1 << 30)
    p.features(True)
    assert p.gpu.get_replay_info() == (0, False)
    p.assert_bitexact(("F", "xf", "dt", "tf"))
    assert np.array_equal(np.array(p.gpu.get_rng_state(), dtype=np.uint64), p.orc.get_rng())


def _replay_seeds():
    from test_gpu_parity_random import NOISY as RNOISY, draw

    seeds = []
    for seed in range(48):
        model, _, observer, *_ = draw(seed)
        if observer in ("thresh2", "nhood2") and model not in RNOISY:
            seeds.append(seed)
    return seeds


@pytest.mark.parametrize("seed", _replay_seeds())
def test_replay_random_configuration_bitexact(gpu_available, seed):
    """The randomly drawn parity configurations with a two-pass observer and
    no noise (every stepper, spans before and across zero, per-member spans,
    event lists), run with the replay on: bit for bit the oracle's."""
    from test_gpu_parity_random import draw

    model, stepper, observer, prec, n, (x0, pars), tspan, sp, per_member, rng = draw(seed)
    print(f"replay seed {seed}: {model} {stepper} {observer} {prec} n={n} tspan={tspan} per_member={per_member}",
          flush=True)
    p = Pair(model, stepper, observer, prec, obs=dict(max_event_timestamps=int(rng.integers(0, 4))), sp=sp)
    p.gpu.set_replay_budget(1 << 30)
    p.set_problem_data(x0, pars)
    p.set_tspan(*tspan)
    if per_member:
        t0 = tspan[0] + rng.uniform(-1.0, 1.0, n)
        p.set_member_tspan(np.vstack([t0, t0 + (tspan[1] - tspan[0]) * rng.uniform(0.2, 1.0, n)]))
    p.transient()
    p.shift_x0()
    p.features()
    assert p.gpu.get_replay_info()[1]
    p.assert_bitexact(("F", "xf", "dt", "tf"))
