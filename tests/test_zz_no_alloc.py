"""clode_mlp_screen_grid_noalloc (the screening entry point that allocates
nothing: caller-owned device outputs, dilation staging in the library's
static device array) against clode_mlp_screen_grid, through the C ABI.

Both run the same kernels on the same weights, so period, amplitude, class
and flags must agree bit for bit.  Two calls in a row, the second on a
smaller grid with another dilation parity, show that what the first call left
in the shared static array does not leak into the second; the output buffers
are pre-filled with junk so that a cell the call fails to write is caught.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_CELLS = 4194304  # CLODE_MLP_NOALLOC_MAX_CELLS, include/clode_mlp.h


def trained_net():
    from clode_amd.surrogate import MLPSurrogate

    m = MLPSurrogate(seed=21)
    rng = np.random.default_rng(8)
    U = rng.random((8192, 2))
    T = np.stack([np.cos(3 * U[:, 1]), U[:, 0] ** 2], axis=1)
    cls = ((U[:, 0] > 0.4).astype(int) + (U[:, 1] > 0.6).astype(int)) * 3
    m.set_target_scaling([1.0, 0.5], [0.5, 2.0])
    m.train(U, T, cls, epochs=4, batch=1024, lr=2e-3, seed=3)
    return m


class HipBuffers:
    """Caller-owned device buffers, allocated through the HIP runtime (the
    caller of the no-allocation entry point allocates its outputs itself)."""

    def __init__(self):
        self.hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def alloc_filled(self, a: np.ndarray) -> int:
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def fetch(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def screen_noalloc(m, nx, ny, tol, dilate):
    from clode_amd.surrogate import _lib

    n = nx * ny
    # junk fills, so that a cell the call fails to write is caught
    init = (np.full(n, np.nan, np.float32), np.full(n, -1.0e30, np.float32), np.full(n, -7, np.int32),
            np.full(n, 0xEE, np.uint8))
    hb = HipBuffers()
    try:
        ptrs = [hb.alloc_filled(a) for a in init]
        lib = _lib()
        rc = lib.clode_mlp_screen_grid_noalloc(m._h, nx, ny, tol, dilate, *ptrs)
        assert rc == 0, lib.clode_mlp_last_error().decode()
        return tuple(hb.fetch(p, a).reshape(nx, ny) for p, a in zip(ptrs, init))
    finally:
        hb.free()


def test_noalloc_screen_matches_allocating_screen(gpu_available):
    m = trained_net()
    for nx, ny, dilate in ((517, 389, 3), (130, 67, 2)):
        print(f"no-alloc screen {nx}x{ny} dilate {dilate}", flush=True)
        ref = m.screen_grid(nx, ny, tol=0.05, dilate=dilate)
        got = screen_noalloc(m, nx, ny, 0.05, dilate)
        for name, a, b in zip(("period", "amp", "cls", "flags"), got, ref):
            assert a.shape == b.shape == (nx, ny)
            assert a.dtype == b.dtype, name
            same = a.view(np.uint8) == b.view(np.uint8)
            assert same.all(), f"{name}: {int((~same).sum())} bytes differ"
        assert np.isfinite(got[0]).all()
        assert (got[3] & 4).any() and (got[3] & 2).any()  # the boundary and its dilation are exercised


def test_noalloc_screen_rejects_oversize_and_host_pointers(gpu_available):
    from clode_amd.surrogate import MLPSurrogate, _lib

    lib = _lib()
    m = MLPSurrogate(seed=1)
    hb = HipBuffers()
    try:
        p = hb.alloc_filled(np.zeros(16, np.float32))
        rc = lib.clode_mlp_screen_grid_noalloc(m._h, 2049, 2048, 0.05, 1, p, p, p, p)
        assert rc == 1
        assert b"limit" in lib.clode_mlp_last_error()
        host = np.zeros(16, np.float32)
        rc = lib.clode_mlp_screen_grid_noalloc(m._h, 4, 4, 0.05, 1, host.ctypes.data, p, p, p)
        assert rc == 1
        assert b"device memory" in lib.clode_mlp_last_error()
    finally:
        hb.free()
    # the handle still works afterwards
    P, A, C, F = m.screen_grid(4, 4, tol=0.05, dilate=1)
    assert P.shape == (4, 4)
    assert MAX_CELLS == 2048 * 2048
