"""The BF16-MFMA surrogate network on the GPU against its CPU oracle, its
training, its screening, and what it gains over the method without it.

Tolerances: the oracle reproduces every BF16 rounding of the device and
accumulates in float64 where the device accumulates in FP32, so outputs agree
to 1e-4 in the median and 1e-2 at worst (where an FP32 sum lands on the other
side of a BF16 rounding boundary, one activation moves by a BF16 ulp and the
change propagates); gradients agree to 1 % in norm per layer with cosine
> 0.9999.  A transposed or misindexed operand gives O(1) errors.
"""
import os

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "clode_amd", "models")


def net(seed=0):
    from clode_amd.surrogate import MLPSurrogate

    return MLPSurrogate(seed=seed)


def test_forward_matches_oracle(gpu_available):
    m = net(3)
    rng = np.random.default_rng(0)
    X = rng.random((1000, 2)).astype(np.float32)
    out, enc = m.forward(X, with_encoding=True)
    # the encoding: sin / cos of the Fourier projections, BF16, within one BF16 ulp
    ref_enc = mo.bf16_round(mo.encode(X.astype(np.float64), m.fourier().astype(np.float64)))
    got_enc = mo.bf16_bits_to_float(enc)
    ulp = 2.0 ** (np.floor(np.log2(np.maximum(np.abs(ref_enc), 1e-30))) - 7)
    assert np.all(np.abs(got_enc - ref_enc) <= ulp * 1.0001)
    assert np.mean(got_enc == ref_enc) > 0.99
    ref, _ = mo.forward(m.get_params().astype(np.float64), got_enc)
    assert np.all(ref[:, mo.NOUT:] == 0)
    err = np.abs(out - ref[:, :mo.NOUT])
    assert np.max(err / (1 + np.abs(ref[:, :mo.NOUT]))) < 1e-2
    assert np.median(err) < 1e-4


def test_forward_large_tiles(gpu_available):
    """Products of 32,768 rows or more use 128 x 128 workgroup tiles; each
    output element still accumulates the same K steps in the same order, so
    the result equals the 64 x 64-tile result bit for bit (40,000 rows at
    once against two halves of 20,000), and the oracle on a sample."""
    m = net(4)
    X = np.random.default_rng(2).random((40000, 2)).astype(np.float32)
    big, enc = m.forward(X, with_encoding=True)
    halves = np.concatenate([m.forward(X[:20000]), m.forward(X[20000:])])
    assert np.array_equal(big, halves)
    idx = np.arange(0, 40000, 97)
    ref, _ = mo.forward(m.get_params().astype(np.float64), mo.bf16_bits_to_float(enc[idx]))
    err = np.abs(big[idx] - ref[:, :mo.NOUT])
    assert np.max(err / (1 + np.abs(ref[:, :mo.NOUT]))) < 1e-2
    assert np.median(err) < 1e-4


def test_null_pointers_are_invalid_arguments(gpu_available):
    """Every entry point refuses a null array or output pointer with
    CLODE_MLP_ERR_INVALID before any device work."""
    from clode_amd.surrogate import _lib

    lib, m = _lib(), net(0)
    h = m._h
    for fn, args in [("clode_mlp_n_params", [None]), ("clode_mlp_get_params", [None, m.n_params]),
                     ("clode_mlp_set_params", [None, m.n_params]), ("clode_mlp_get_fourier", [None, 32]),
                     ("clode_mlp_set_target_scaling", [None, None]),
                     ("clode_mlp_forward", [None, 10, None, None]),
                     ("clode_mlp_loss_grad", [None, None, None, 10, None, None]),
                     ("clode_mlp_train", [None, None, None, 10, 1, 64, 1e-3, 0, None]),
                     ("clode_mlp_screen_grid", [4, 4, 0.1, 0, None, None, None, None]),
                     ("clode_mlp_last_ms", [None])]:
        assert getattr(lib, fn)(h, *args) == 1, fn
        assert b"null" in lib.clode_mlp_last_error(), fn
    # the handle still works afterwards
    assert m.forward(np.zeros((3, 2), np.float32)).shape[0] == 3


def test_gradient_matches_oracle(gpu_available):
    m = net(5)
    rng = np.random.default_rng(1)
    n = 3000  # not a multiple of the 64-row tile
    X = rng.random((n, 2)).astype(np.float32)
    T = rng.normal(0, 1, (n, 2)).astype(np.float32)
    cls = rng.integers(0, 16, n)
    loss, g = m.loss_grad(X, T, cls)
    _, enc = m.forward(X, with_encoding=True)
    rl, rg = mo.loss_grad(m.get_params().astype(np.float64), mo.bf16_bits_to_float(enc), T.astype(np.float64), cls)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    off = 0
    for fin, fout in mo.LAYERS:
        for part in (fin * fout, fout):
            a, b = g[off:off + part].astype(np.float64), rg[off:off + part]
            off += part
            nb = np.linalg.norm(b)
            assert np.linalg.norm(a - b) <= 1e-2 * nb + 1e-12, (fin, fout, part)
            if nb > 0:
                assert a @ b / (np.linalg.norm(a) * nb) > 0.9999
    assert off == g.size


def smooth_targets(U):
    lp = np.sin(3 * U[:, 0]) + 0.5 * np.cos(2 * U[:, 1])
    amp = U[:, 0] * U[:, 1] - 0.3
    cls = np.clip((4 * U[:, 0] + 2 * U[:, 1]).astype(int), 0, 15)
    return np.stack([lp, amp], axis=1), cls


def test_training_fits_a_smooth_map(gpu_available):
    m = net(7)
    rng = np.random.default_rng(2)
    U = rng.random((16384, 2))
    T, cls = smooth_targets(U)
    hist = m.train(U, T, cls, epochs=30, batch=1024, lr=2e-3, seed=1)
    assert np.all(np.isfinite(hist)) and hist[-1] < hist[0] / 10
    V = rng.random((4000, 2))
    Tv, cv = smooth_targets(V)
    out = m.forward(V)
    for h in range(4):
        z = out[:, h * mo.HW:(h + 1) * mo.HW]
        assert np.median(np.abs(z[:, :2] - Tv)) < 0.03
        assert np.mean(np.argmax(z[:, 2:], axis=1) == cv) > 0.95


@pytest.mark.parametrize("n, batch", [(16384, 1024), (5000, 2048), (3000, 64)])
def test_fused_optimizer_matches_separate_sums(gpu_available, n, batch):
    """The training loop's fused launches (each layer's weight and input
    gradients in one launch, mlp_gemm_pair; the split-K sums and bias
    reductions inside the optimizer launch, mlp_adam_sum) against every
    product in a launch of its own, the backward pass's reduction launches
    and a plain Adam launch (CLODE_MLP_FUSED_TRAIN=0): same blocks, partial
    products, summation order and tree, so the trained parameters agree bit
    for bit (the reported epoch losses are summed with atomics, in no fixed
    order, so only to rounding).  Full and ragged last minibatches, and
    minibatches too small to split."""
    rng = np.random.default_rng(8)
    U = rng.random((n, 2))
    T, cls = smooth_targets(U)
    fused = net(13)
    h_fused = fused.train(U, T, cls, epochs=3, batch=batch, lr=2e-3, seed=6)
    os.environ["CLODE_MLP_FUSED_TRAIN"] = "0"
    try:
        separate = net(13)
        h_sep = separate.train(U, T, cls, epochs=3, batch=batch, lr=2e-3, seed=6)
    finally:
        del os.environ["CLODE_MLP_FUSED_TRAIN"]
    assert np.array_equal(fused.get_params(), separate.get_params())
    assert np.allclose(h_fused, h_sep, rtol=1e-5, atol=0)
    assert not np.array_equal(fused.get_params(), net(13).get_params())  # training moved the parameters


def test_screen_flags_class_boundary(gpu_available):
    m = net(9)
    rng = np.random.default_rng(3)
    U = rng.random((16384, 2))
    T = np.zeros((U.shape[0], 2))
    cls = (U[:, 0] > 0.5).astype(int) * 3
    m.set_target_scaling([0.0, 1.0], [1.0, 1.0])
    m.train(U, T, cls, epochs=30, batch=1024, lr=2e-3, seed=4)
    P, A, C, F = m.screen_grid(101, 37, tol=0.05, dilate=1)
    assert P.shape == (101, 37)
    assert np.all(C[:45] == 0) and np.all(C[56:] == 3)
    # the boundary (x = 0.5, cell 50) is flagged as a class change in every
    # column, and nowhere far from it
    assert np.all((F[48:53] & 2).max(axis=0) > 0)
    assert np.all((F[:40] & 2) == 0) and np.all((F[61:] & 2) == 0)
    assert set(np.unique(F)) <= {0, 1, 2, 3, 4, 5, 6, 7}


@pytest.mark.parametrize("shape", [(517, 389), (600, 500), (1, 1), (3, 129)])
def test_screen_fused_matches_layer_by_layer(gpu_available, shape):
    """The fused screening kernel (whole network per 128-cell block, the
    activations in LDS) against the layer-by-layer products over chunks of
    2^18 cells (CLODE_MLP_FUSED_SCREEN=0): same operand layout, K order and
    epilogues, so period, amplitude, class and flags agree bit for bit.
    Grids that are not multiples of the block, one that crosses the chunk
    boundary (300,000 cells), a single cell and a one-row grid."""
    m = net(11)
    rng = np.random.default_rng(5)
    U = rng.random((8192, 2))
    T = np.stack([np.sin(4 * U[:, 0]), U[:, 1] ** 2], axis=1)
    cls = (U[:, 0] + U[:, 1] > 1.0).astype(int) * 2
    m.set_target_scaling([1.0, 0.5], [0.5, 2.0])
    m.train(U, T, cls, epochs=4, batch=1024, lr=2e-3, seed=2)
    nx, ny = shape
    fused = m.screen_grid(nx, ny, tol=0.05, dilate=1)
    os.environ["CLODE_MLP_FUSED_SCREEN"] = "0"
    try:
        layered = m.screen_grid(nx, ny, tol=0.05, dilate=1)
    finally:
        del os.environ["CLODE_MLP_FUSED_SCREEN"]
    for a, b in zip(fused, layered):
        assert a.shape == (nx, ny)
        assert np.array_equal(a, b)
    assert np.all(np.isfinite(fused[0])) and np.all(np.isfinite(fused[1]))


def lactotroph_solver(n_axis_check=None):
    import clode_amd as clode
    from clode_amd.surrogate import integrator_solver

    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph.cl"),
                                 variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                 parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0}, aux=["ica"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                 single_precision="fp32c", dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, observer_x_up_thresh=0.3, observer_x_down_thresh=0.2)
    return integrator_solver(sim, ("gk", "gbk"), ((0.0, 10.0), (0.0, 2.0)), 1000.0, 6000.0, "v")


def test_screened_map_and_ablation(gpu_available):
    """On a 128 x 128 lactotroph map (fp32c): the trained surrogate re-solves
    a minority of cells and the final map disagrees with the plain map on at
    most 1 % of cells; an untrained network (the method without the learned
    part) has to re-solve nearly everything to reach the same map."""
    from clode_amd.surrogate import compare_maps, screened_map

    solve = lactotroph_solver()
    nx = ny = 128
    gx, gy = np.meshgrid(np.arange(nx) / (nx - 1), np.arange(ny) / (ny - 1), indexing="ij")
    plain = solve(np.stack([gx.ravel(), gy.ravel()], axis=1))
    plain = {k: v.reshape(nx, ny) for k, v in plain.items()}
    m, net_ = screened_map(solve, nx, ny, n_train=8192, epochs=60, batch=512, lr=2e-3, tol=0.05, dilate=1,
                           fix_rounds=3, seed=1)
    bad = compare_maps({"period": m.period, "amp": m.amp, "cls": m.cls}, plain, tol=0.05)
    frac_solved = m.solved.mean()
    assert bad.mean() <= 0.01, (bad.mean(), frac_solved)
    assert frac_solved < 0.6, frac_solved
    # without training: heads disagree almost everywhere
    from clode_amd.surrogate import MLPSurrogate

    raw = MLPSurrogate(seed=11)
    raw.set_target_scaling(net_.target_mean, net_.target_std)
    P, A, C, F = raw.screen_grid(nx, ny, tol=0.05, dilate=1)
    untrained_flagged = (F != 0).mean()
    unflagged = F == 0
    untrained_bad = compare_maps({"period": P.astype(float), "amp": A.astype(float), "cls": C}, plain)[unflagged]
    assert untrained_flagged > 0.9 or untrained_bad.mean() > 0.2
