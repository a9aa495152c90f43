"""Work-queue launches, member time spans and batches of different systems,
against the CPU oracle and against the plain grid launch.

* a simulator in work-queue mode gives bit-identical results to the oracle
  (plain single and double precision) with uneven per-member time spans;
* a batch of four different systems (Van der Pol rk4 + thresh2, Lorenz
  dopri5 + localmax, Chay-Keizer dopri5 + thresh2, Ornstein-Uhlenbeck seuler
  + basicall), each with lognormal per-member spans, matches four separate
  oracle runs bit for bit;
* fp32c (no bit-exact oracle): batch results equal the same simulators run
  one by one with grid launches, and a graph-captured sweep equals the eager
  sweep;
* the autotuner leaves every member's state as it found it, times every
  launch shape three times and keeps the fastest median, and a later batch of
  the same program and sizes takes the stored choice without trials.
"""
import numpy as np
import pytest

from helpers import Pair

pytestmark = pytest.mark.gpu

ADAPT = (0.1, 1.0, 1e-8, 1e-6, 10 ** 7, 1000000, 1)
FIXED = (0.01, 1.0, 1e-6, 1e-3, 10 ** 7, 1000000, 1)
NOISE = (0.05, 1.0, 1e-6, 1e-3, 10 ** 7, 1000000, 1)


def lognormal_spans(n, median, sigma, seed):
    rng = np.random.default_rng(seed)
    t1 = median * np.exp(sigma * rng.standard_normal(n))
    return np.vstack([np.zeros(n), t1])


def make_system(name, prec, n, seed, oracle=True):
    """One of the four systems of the mixed batch, with its data and spans."""
    rng = np.random.default_rng(seed)
    if name == "vdp":
        p = Pair("van_der_pol", "rk4", "thresh2", prec, sp=FIXED, oracle=oracle)
        p.set_problem_data(np.vstack([2.0 * np.ones(n), np.zeros(n)]), np.linspace(0.1, 5.0, n)[None, :])
        spans = lognormal_spans(n, 40.0, 0.5, seed)
    elif name == "lorenz":
        p = Pair("lorenz", "dopri5", "localmax", prec, sp=ADAPT, obs=dict(max_event_timestamps=2), oracle=oracle)
        p.set_problem_data(rng.normal(1.0, 0.1, (3, n)),
                           np.vstack([10.0 * np.ones(n), np.linspace(0, 30, n), np.full(n, 8 / 3)]))
        spans = lognormal_spans(n, 15.0, 0.7, seed)
    elif name == "ck":
        p = Pair("chay_keizer", "dopri5", "thresh2", prec, sp=ADAPT, oracle=oracle)
        x0 = np.tile(np.array([[-50.0], [0.01], [0.12]]), (1, n))
        pars = np.vstack([np.linspace(550, 1050, n), 750.0 * np.ones(n), np.linspace(0.095, 0.155, n)])
        p.set_problem_data(x0, pars)
        spans = lognormal_spans(n, 300.0, 0.6, seed)
    else:
        p = Pair("ornstein_uhlenbeck", "seuler", "basicall", prec, sp=NOISE, oracle=oracle)
        p.set_problem_data(np.zeros((1, n)), np.vstack([np.ones(n), np.ones(n), 0.5 * np.ones(n)]))
        spans = lognormal_spans(n, 20.0, 1.0, seed)
    p.set_tspan(0.0, 1.0)  # replaced member by member
    p.set_member_tspan(spans)
    p.seed(1000 + seed)
    return p


def batch_of(pairs):
    from clode_amd.cpp import clode_cpp_wrapper as w

    return w.BatchSimulatorBase([p.gpu for p in pairs])


@pytest.mark.parametrize("prec", ["double", "single"])
def test_member_tspan_grid_and_queue_bitexact(gpu_available, prec):
    n = 150
    for mode in (0, 1):
        p = Pair("lactotroph", "dopri5", "thresh2", prec, sp=ADAPT, obs=dict(max_event_timestamps=2))
        g = np.linspace(0.0, 10.0, n)
        x0 = np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n))
        p.set_problem_data(x0, np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), np.linspace(2, 0, n), g]))
        p.set_tspan(0.0, 1.0)
        p.set_member_tspan(lognormal_spans(n, 400.0, 0.8, 11))
        p.gpu.set_launch(mode, 64 if mode == 0 else 128)
        p.transient()
        p.assert_bitexact(("xf", "dt", "tf"))
        p.shift_x0()
        p.features()
        p.assert_bitexact(("F", "xf", "dt", "tf"))
        tf = np.array(p.gpu.get_tf())
        t1 = lognormal_spans(n, 400.0, 0.8, 11)[1]
        assert np.all(tf >= t1 * (1 - 1e-6)) and np.all(tf <= t1 + 2.0)  # each stopped at its own end


@pytest.mark.parametrize("prec", ["double", "single"])
def test_queue_trajectory_bitexact(gpu_available, prec):
    n = 77
    p = Pair("van_der_pol", "dopri5", "", prec, kind="trajectory", sp=(0.05, 1.0, 1e-8, 1e-6, 10 ** 7, 300, 2))
    p.set_problem_data(np.vstack([2.0 * np.ones(n), np.zeros(n)]), np.linspace(0.1, 5.0, n)[None, :])
    p.set_tspan(0.0, 1.0)
    p.set_member_tspan(lognormal_spans(n, 30.0, 0.7, 5))
    p.gpu.set_launch(1, 64)
    p.gpu.set_queue_slice(3)
    p.trajectory()
    assert p.assert_trajectory_bitexact() > n


@pytest.mark.parametrize("prec", ["double", "single"])
def test_mixed_batch_matches_separate_oracles(gpu_available, prec):
    names = ("vdp", "lorenz", "ck", "ou")
    sizes = (130, 97, 64, 201)
    pairs = [make_system(nm, prec, n, k) for k, (nm, n) in enumerate(zip(names, sizes))]
    b = batch_of(pairs)
    b.set_launch(128, 4)
    b.transient()
    for p in pairs:
        p.orc.transient()
        p.assert_bitexact(("xf", "dt", "tf"))
    b.shift_x0()
    for p in pairs:
        p.orc.shift_x0()
    b.features()
    for p in pairs:
        p.orc.features()
        res = p.assert_bitexact(("F", "xf", "dt", "tf"))
        assert res["F"][1] == p.n * p.gpu.get_n_features()
    # continuation: a second features() continues every observer
    b.shift_x0()
    b.features()
    for p in pairs:
        p.orc.shift_x0()
        p.orc.features()
        p.assert_bitexact(("F", "xf"))


def run_fp32c_systems(launch):
    """The four systems in fp32c with identical data; 'grid' runs them one by
    one, 'eager' / 'graph' as one batch sweep of two rounds."""
    pairs = [make_system(nm, "fp32c", n, k, oracle=False)
             for k, (nm, n) in enumerate(zip(("vdp", "lorenz", "ck", "ou"), (300, 257, 128, 513)))]
    if launch == "grid":
        for _ in range(2):
            for p in pairs:
                p.gpu.transient()
                p.gpu.shift_x0()
                p.gpu.features()
                p.gpu.shift_x0()
    else:
        b = batch_of(pairs)
        b.set_launch(64, 16)
        b.sweep(2, launch == "graph")
    return [(np.array(p.gpu.get_f()), np.array(p.gpu.get_x0()), np.array(p.gpu.get_dt())) for p in pairs]


def test_fp32c_batch_sweep_graph_eager_and_grid_agree(gpu_available):
    ref = run_fp32c_systems("grid")
    for launch in ("eager", "graph"):
        out = run_fp32c_systems(launch)
        for (f, x, d), (fr, xr, dr) in zip(out, ref):
            assert np.array_equal(f, fr) and np.array_equal(x, xr) and np.array_equal(d, dr), launch


def test_graph_sweep_twice_equals_eager_four_rounds(gpu_available):
    pairs = [make_system(nm, "fp32c", n, k, oracle=False)
             for k, (nm, n) in enumerate(zip(("vdp", "ou"), (200, 300)))]
    b = batch_of(pairs)
    b.sweep(2, True)
    b.sweep(2, True)
    got = [np.array(p.gpu.get_f()) for p in pairs]
    pairs2 = [make_system(nm, "fp32c", n, k, oracle=False)
              for k, (nm, n) in enumerate(zip(("vdp", "ou"), (200, 300)))]
    b2 = batch_of(pairs2)
    b2.sweep(4, False)
    for a, r in zip(got, [np.array(p.gpu.get_f()) for p in pairs2]):
        assert np.array_equal(a, r)


def test_autotune_restores_state(gpu_available, monkeypatch):
    pairs = [make_system(nm, "double", n, k) for k, (nm, n) in enumerate(zip(("lorenz", "ou"), (300, 400)))]
    b = batch_of(pairs)
    monkeypatch.setenv("CLODE_AUTOTUNE_FRESH", "1")
    block, slice_ = b.autotune(4)
    assert block in (64, 128, 256) and slice_ in (1, 4, 16, 64)
    assert tuple(b.get_launch()) == (block, slice_)
    rows, cached = b.autotune_table()
    rows = np.asarray(rows)
    assert not cached and rows.shape == (12, 5)
    assert {(int(r[0]), int(r[1])) for r in rows} == {(bb, ss) for bb in (64, 128, 256) for ss in (1, 4, 16, 64)}
    assert np.all(rows[:, 3] <= rows[:, 2]) and np.all(rows[:, 2] <= rows[:, 4]) and np.all(rows[:, 3] > 0)
    best = rows[np.argmin(rows[:, 2])]
    assert (int(best[0]), int(best[1])) == (block, slice_)
    b.transient()
    for p in pairs:
        p.orc.transient()
        p.assert_bitexact(("xf", "dt", "tf"))
    # the choice persists: the same program and sizes take it without trials
    monkeypatch.delenv("CLODE_AUTOTUNE_FRESH")
    b2 = batch_of([make_system(nm, "double", n, k, oracle=False)
                   for k, (nm, n) in enumerate(zip(("lorenz", "ou"), (300, 400)))])
    b2.set_launch(256, 64 if (block, slice_) != (256, 64) else 1)
    assert tuple(b2.autotune(4)) == (block, slice_)
    rows2, cached2 = b2.autotune_table()
    assert cached2 and np.asarray(rows2).shape == (0, 5)
    assert tuple(b2.get_launch()) == (block, slice_)


def test_batch_validation(gpu_available):
    from clode_amd.cpp import clode_cpp_wrapper as w

    a = make_system("vdp", "double", 10, 0, oracle=False)
    c = make_system("ou", "single", 10, 1, oracle=False)
    with pytest.raises(ValueError):
        w.BatchSimulatorBase([a.gpu, c.gpu]).transient()  # mixed precision
    with pytest.raises(ValueError):
        w.BatchSimulatorBase([a.gpu, a.gpu])
    with pytest.raises(ValueError):
        w.BatchSimulatorBase([])
