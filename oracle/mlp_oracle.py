"""CPU restatement of the feature-map surrogate (clode_amd/csrc/mlp/clode_mlp.hip).

Test infrastructure only: used by tests/ to check the GPU network's forward
pass, loss and parameter gradients; never by the product path.

The restatement follows the device code step by step: BF16 weights and
activations (round to nearest even at exactly the points where the device
stores BF16), FP32-or-better accumulation (float64 here), SiLU, the per-head
bootstrap mask, the loss (mean over rows of the sum over heads of two squared
errors and a softmax cross-entropy) and its gradient.  With ``bf16=False``
every rounding is skipped and the result is the exact float64 function, whose
analytic gradient the tests pin against central finite differences.
"""
from __future__ import annotations

import numpy as np

NFF, D0, H, NHID, NHEAD, NCLS = 16, 32, 256, 4, 4, 16
HW = 2 + NCLS
NOUT = NHEAD * HW
NOUTP = 128
LAYERS = [(D0, H)] + [(H, H)] * (NHID - 1) + [(H, NOUTP)]


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float -> bf16 -> float, round to nearest even (as the device f2bf)."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def bf16_bits_to_float(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray):
    """Flat vector -> [(Wt[out][in], b[out])] in layer order."""
    out, off = [], 0
    for fin, fout in LAYERS:
        W = params[off:off + fin * fout].reshape(fout, fin)
        off += fin * fout
        b = params[off:off + fout]
        off += fout
        out.append((W, b))
    assert off == params.size
    return out


def encode(X: np.ndarray, B: np.ndarray) -> np.ndarray:
    """[n][2] -> [n][32]: sin, cos of 2 pi (B[0][f] x0 + B[1][f] x1), interleaved."""
    p = 2 * np.pi * (X[:, :1] * B[0][None, :] + X[:, 1:2] * B[1][None, :])
    E = np.empty((X.shape[0], D0))
    E[:, 0::2] = np.sin(p)
    E[:, 1::2] = np.cos(p)
    return E


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def head_keeps(ids: np.ndarray, head: int) -> np.ndarray:
    """Bootstrap membership (splitmix64 finaliser of 4 * id + head + golden)."""
    with np.errstate(over="ignore"):
        z = ids.astype(np.uint64) * np.uint64(4) + np.uint64(head) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return (z & np.uint64(3)) != 0


def forward(params: np.ndarray, A0: np.ndarray, bf16: bool = True):
    """Head outputs [n][128] and the cache for backward.  A0 is the encoded
    input (already BF16-valued when bf16=True)."""
    rnd = bf16_round if bf16 else (lambda v: v)
    layers = unpack(np.asarray(params, dtype=np.float64))
    Ys, Zs = [A0], []
    y = A0
    for l, (W, b) in enumerate(layers):
        z = y @ rnd(W).T + b
        Zs.append(z)
        if l < NHID:
            y = rnd(silu(z))
            Ys.append(y)
    return Zs[-1], (layers, Ys, Zs)


def loss_and_dout(Zh: np.ndarray, T: np.ndarray, cls: np.ndarray, ids: np.ndarray):
    n = Zh.shape[0]
    loss = 0.0
    d = np.zeros_like(Zh)
    for h in range(NHEAD):
        keep = head_keeps(ids, h)
        zh = Zh[:, h * HW:(h + 1) * HW]
        e = zh[:, :2] - T
        logits = zh[:, 2:]
        mx = logits.max(axis=1, keepdims=True)
        ex = np.exp(logits - mx)
        se = ex.sum(axis=1, keepdims=True)
        ce = (mx[:, 0] + np.log(se[:, 0])) - logits[np.arange(n), cls]
        loss += np.sum(((e ** 2).sum(axis=1) + ce)[keep])
        g = np.zeros((n, HW))
        g[:, :2] = 2 * e
        p = ex / se
        p[np.arange(n), cls] -= 1.0
        g[:, 2:] = p
        g[~keep] = 0
        d[:, h * HW:(h + 1) * HW] = g / n
    return loss / n, d


def loss_grad(params: np.ndarray, A0: np.ndarray, T: np.ndarray, cls: np.ndarray, ids: np.ndarray | None = None,
              bf16: bool = True):
    """Loss and flat gradient; the device's BF16 roundings of the output
    gradient and of each back-propagated activation gradient are applied when
    bf16=True."""
    rnd = bf16_round if bf16 else (lambda v: v)
    if ids is None:
        ids = np.arange(A0.shape[0], dtype=np.uint64)
    Zh, (layers, Ys, Zs) = forward(params, A0, bf16)
    loss, dz = loss_and_dout(Zh, T, np.asarray(cls), ids)
    dz = rnd(dz)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (dz.T @ Ys[l], dz.sum(axis=0))
        if l == 0:
            break
        dz = rnd((dz @ rnd(W)) * dsilu(Zs[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat
