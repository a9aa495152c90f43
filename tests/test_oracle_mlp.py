"""Pins the surrogate's CPU oracle (oracle/mlp_oracle.py):
  * BF16 rounding against its definition (nearest, ties to even);
  * the analytic loss gradient against central finite differences of the
    exact float64 network (no BF16 rounding), parameter by parameter, for
    weights and biases of every layer, regression and class outputs;
  * the bootstrap mask keeps about 3/4 of the samples per head, differently
    per head.
"""
import numpy as np
import pytest

import mlp_oracle as mo


def test_bf16_round_nearest_even():
    # 1 + 2^-8 is halfway between bf16 1.0 and 1 + 2^-7: ties to even -> 1.0
    assert mo.bf16_round(np.array([1 + 2 ** -8]))[0] == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert mo.bf16_round(np.array([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6
    assert mo.bf16_round(np.array([1 + 2 ** -8 + 2 ** -20]))[0] == 1 + 2 ** -7
    rng = np.random.default_rng(0)
    x = rng.normal(0, 10, 10000).astype(np.float32)
    r = mo.bf16_round(x)
    # representable in 8 significant bits and within half a bf16 ulp
    m, e = np.frexp(r)
    assert np.all(m * 256 == np.rint(m * 256))
    ulp = 2.0 ** (np.floor(np.log2(np.abs(x.astype(np.float64)))) - 7)
    assert np.all(np.abs(r - x) <= ulp / 2 * (1 + 1e-12))


def random_problem(n=12, seed=1):
    rng = np.random.default_rng(seed)
    sizes = sum(fin * fout + fout for fin, fout in mo.LAYERS)
    p = np.zeros(sizes)
    off = 0
    for l, (fin, fout) in enumerate(mo.LAYERS):
        W = rng.normal(0, np.sqrt(2.0 / fin), (fout, fin))
        if l == len(mo.LAYERS) - 1:
            W[mo.NOUT:] = 0
        p[off:off + fin * fout] = W.ravel()
        off += fin * fout
        b = rng.normal(0, 0.1, fout)
        if l == len(mo.LAYERS) - 1:
            b[mo.NOUT:] = 0
        p[off:off + fout] = b
        off += fout
    X = rng.random((n, 2))
    A0 = mo.encode(X, rng.normal(0, 3, (2, 16)))
    T = rng.normal(0, 1, (n, 2))
    cls = rng.integers(0, 16, n)
    return p, A0, T, cls


def test_analytic_gradient_matches_finite_differences():
    p, A0, T, cls = random_problem()
    loss, g = mo.loss_grad(p, A0, T, cls, bf16=False)
    rng = np.random.default_rng(2)
    # indices in every layer's weights and biases, including the head's
    # regression and class columns
    idx, off = [], 0
    for l, (fin, fout) in enumerate(mo.LAYERS):
        nw = fin * fout
        rows = rng.integers(0, mo.NOUT if l == len(mo.LAYERS) - 1 else fout, 6)
        cols = rng.integers(0, fin, 6)
        idx += [off + r * fin + c for r, c in zip(rows, cols)]
        idx += [off + nw + int(r) for r in rows[:3]]
        off += nw + fout
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp, _ = mo.loss_grad(q, A0, T, cls, bf16=False)
        q[k] -= 2 * h
        lm, _ = mo.loss_grad(q, A0, T, cls, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(fd)), (k, fd, g[k])


def test_padding_columns_get_no_gradient():
    p, A0, T, cls = random_problem(n=8, seed=3)
    _, g = mo.loss_grad(p, A0, T, cls, bf16=True)
    layers = mo.unpack(g)
    W, b = layers[-1]
    assert np.all(W[mo.NOUT:] == 0) and np.all(b[mo.NOUT:] == 0)
    assert np.abs(W[:mo.NOUT]).max() > 0


def test_head_bootstrap_masks():
    ids = np.arange(40000, dtype=np.uint64)
    keeps = [mo.head_keeps(ids, h) for h in range(4)]
    for k in keeps:
        assert abs(k.mean() - 0.75) < 0.01
    assert not np.array_equal(keeps[0], keeps[1])
