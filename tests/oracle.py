"""ctypes front end of the CPU oracle (oracle/oracle_model.cpp).

Test infrastructure: used by the tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg to verify GPU results.  ``OracleEnsemble`` mirrors the reference's
host API (set_problem_data, set_tspan, set_solver_params, seed_rng,
transient, features, trajectory, shift_x0, getters), so a test drives the GPU
simulator and the oracle with the same call sequence.
"""
from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "oracle"))
import build_oracle  # noqa: E402

STEPPERS = {"euler": 1, "heun": 2, "rk4": 3, "bs23": 4, "dopri5": 5, "seuler": 6, "he12": 7}
OBSERVERS = {"": 0, "thresh2": 1, "localmax": 2, "nhood2": 3, "basic": 4, "basicall": 5, "nhood1": 6}

_libs: dict[str, ctypes.CDLL] = {}


def _lib(model: str, variant: str = "") -> ctypes.CDLL:
    key = model + ("/" + variant if variant else "")
    if key not in _libs:
        path = build_oracle.oracle_path(model, variant)
        if variant or not path.exists():
            path = build_oracle.build_model(model, variant=variant)
        lib = ctypes.CDLL(str(path))
        vp, sz, dp, ip = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        u64p = ctypes.POINTER(ctypes.c_uint64)
        lib.oracle_create.restype = vp
        lib.oracle_create.argtypes = [ip] * 9
        lib.oracle_destroy.argtypes = [vp]
        lib.oracle_set_threads.argtypes = [vp, ip]
        lib.oracle_set_problem_data.argtypes = [vp, dp, sz, dp, sz]
        lib.oracle_set_x0.argtypes = [vp, dp, sz]
        lib.oracle_set_tspan.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        lib.oracle_set_solver_params.argtypes = [vp] + [ctypes.c_double] * 4 + [ctypes.c_uint] * 3 + [ip]
        lib.oracle_set_observer_params.argtypes = [vp, ctypes.POINTER(ctypes.c_uint), dp]
        lib.oracle_seed_rng.argtypes = [vp, ip]
        lib.oracle_set_rng.argtypes = [vp, u64p, sz]
        lib.oracle_set_dt.argtypes = [vp, dp, sz]
        lib.oracle_set_member_tspan.argtypes = [vp, dp, sz]
        for f in ("oracle_transient", "oracle_initialize_observer", "oracle_trajectory", "oracle_shift_x0"):
            getattr(lib, f).argtypes = [vp]
        lib.oracle_features.argtypes = [vp, ip]
        lib.oracle_last_steps.restype = ctypes.c_ulonglong
        lib.oracle_last_steps.argtypes = [vp]
        lib.oracle_last_trials.restype = ctypes.c_ulonglong
        lib.oracle_last_trials.argtypes = [vp]
        lib.oracle_last_warm_steps.restype = ctypes.c_ulonglong
        lib.oracle_last_warm_steps.argtypes = [vp]
        lib.oracle_size.restype = ctypes.c_longlong
        lib.oracle_size.argtypes = [vp, ctypes.c_char_p]
        lib.oracle_get.argtypes = [vp, ctypes.c_char_p, dp, sz]
        lib.oracle_get_rng.argtypes = [vp, u64p, sz]
        lib.oracle_xoroshiro_next.argtypes = [u64p, u64p, sz]
        lib.oracle_randn.argtypes = [ip, u64p, dp, sz]
        lib.oracle_cm_eval.argtypes = [ip, ip, dp, dp, dp, sz]
        if variant == "count":
            lib.oracle_counts.argtypes = [dp, sz]
        _libs[key] = lib
    return _libs[key]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u64ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class OracleEnsemble:
    """Stateful CPU restatement of one simulator (base, features or trajectory)."""

    def __init__(self, model: str, n_var: int, n_par: int, n_aux: int = 0, n_noise: int = 0,
                 stepper: str = "rk4", observer: str = "", precision: str = "double",
                 n_store: int = 0, kind: str = "features", threads: int | None = None, variant: str = ""):
        if precision == "count":  # FP64 semantics on the operation-counting type
            variant = "count"
        self.lib = _lib(model, variant)
        prec = {"single": 1, "double": 2, "count": 9}[precision]
        kinds = {"base": 0, "features": 1, "trajectory": 2}
        self.h = self.lib.oracle_create(kinds[kind], prec, STEPPERS[stepper], OBSERVERS[observer],
                                        n_var, n_par, n_aux, n_noise, n_store)
        self.n_var, self.n_par = n_var, n_par
        self.lib.oracle_set_threads(self.h, threads if threads else (os.cpu_count() or 1))
        self.refresh_dt = 0

    def __del__(self):
        try:
            self.lib.oracle_destroy(self.h)
        except Exception:
            pass

    # ---- setters (host semantics of clode/cpp/CLODE.cpp)
    def set_problem_data(self, x0, pars):
        x0 = np.ascontiguousarray(x0, dtype=np.float64).ravel()
        pars = np.ascontiguousarray(pars, dtype=np.float64).ravel()
        rc = self.lib.oracle_set_problem_data(self.h, _dptr(x0), x0.size, _dptr(pars), pars.size)
        assert rc == 0, "oracle rejected problem data"

    def set_x0(self, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64).ravel()
        assert self.lib.oracle_set_x0(self.h, _dptr(x0), x0.size) == 0

    def set_tspan(self, tspan):
        self.lib.oracle_set_tspan(self.h, float(tspan[0]), float(tspan[1]))

    def set_member_tspan(self, tspan_m):
        """Per-member spans [2][nPts] (row 0 start, row 1 end); None clears."""
        if tspan_m is None:
            assert self.lib.oracle_set_member_tspan(self.h, None, 0) == 0
            return
        t = np.ascontiguousarray(tspan_m, dtype=np.float64).ravel()
        assert self.lib.oracle_set_member_tspan(self.h, _dptr(t), t.size) == 0

    def set_solver_params(self, dt=0.1, dtmax=0.5, abstol=1e-6, reltol=1e-3, max_steps=1000000,
                          max_store=1000000, nout=1):
        self.lib.oracle_set_solver_params(self.h, dt, dtmax, abstol, reltol, max_steps, max_store, nout,
                                          self.refresh_dt)

    def set_observer_params(self, e_var_ix=0, f_var_ix=0, max_event_count=100, max_event_timestamps=0,
                            min_amp=0.0, min_imi=0.0, nhood_radius=0.05, x_up_threshold=0.2,
                            x_down_threshold=0.2, dx_up_threshold=0.0, dx_down_threshold=0.0, eps_dx=0.0):
        u = (ctypes.c_uint * 4)(e_var_ix, f_var_ix, max_event_count, max_event_timestamps)
        r = np.array([min_amp, min_imi, nhood_radius, x_up_threshold, x_down_threshold, dx_up_threshold,
                      dx_down_threshold, eps_dx], dtype=np.float64)
        self.lib.oracle_set_observer_params(self.h, u, _dptr(r))

    def seed_rng(self, seed: int):
        self.lib.oracle_seed_rng(self.h, int(seed))

    def set_rng(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        assert self.lib.oracle_set_rng(self.h, _u64ptr(w), w.size) == 0

    def set_dt(self, dt):
        d = np.ascontiguousarray(dt, dtype=np.float64)
        assert self.lib.oracle_set_dt(self.h, _dptr(d), d.size) == 0

    # ---- compute
    def transient(self):
        self.lib.oracle_transient(self.h)

    def initialize_observer(self):
        self.lib.oracle_initialize_observer(self.h)

    def features(self, reinit: int = -1):
        self.lib.oracle_features(self.h, reinit)

    def trajectory(self):
        self.lib.oracle_trajectory(self.h)

    def shift_x0(self):
        self.lib.oracle_shift_x0(self.h)

    @property
    def last_steps(self) -> int:
        return int(self.lib.oracle_last_steps(self.h))

    @property
    def last_trials(self) -> int:
        """Step trials of the last pass: its steps plus the rejected trials."""
        return int(self.lib.oracle_last_trials(self.h))

    @property
    def last_warm_steps(self) -> int:
        return int(self.lib.oracle_last_warm_steps(self.h))

    # ---- getters (flat, reference layouts)
    def get(self, name: str) -> np.ndarray:
        n = self.lib.oracle_size(self.h, name.encode())
        out = np.zeros(max(n, 0), dtype=np.float64)
        assert self.lib.oracle_get(self.h, name.encode(), _dptr(out), out.size) == 0
        return out

    COUNT_NAMES = ("add", "mul", "div", "fma", "exp", "log", "pow", "pown", "sqrt", "other")
    PHASES = ("transient", "warmup", "features", "trajectory")
    BUCKETS = ("rejected", "accepted", "failed", "all")

    def counts(self) -> dict:
        """Operation counts of an ensemble created with precision="count":
        {phase: {"trials": {accepted, rejected, failed},
                 bucket: {count name: n}}} (oracle/flop_count.hpp)."""
        nc, npb = len(self.COUNT_NAMES), len(self.BUCKETS)
        out = np.zeros(len(self.PHASES) * npb * nc + len(self.PHASES) * 3)
        assert self.lib.oracle_counts(_dptr(out), out.size) == 0
        res = {}
        for p, ph in enumerate(self.PHASES):
            d = {}
            for b, bk in enumerate(self.BUCKETS):
                base = (p * npb + b) * nc
                d[bk] = {nm: int(out[base + c]) for c, nm in enumerate(self.COUNT_NAMES)}
            t0 = len(self.PHASES) * npb * nc + 3 * p
            d["trials"] = {"rejected": int(out[t0]), "accepted": int(out[t0 + 1]), "failed": int(out[t0 + 2])}
            res[ph] = d
        return res

    def get_rng(self) -> np.ndarray:
        n = self.lib.oracle_size(self.h, b"rng")
        out = np.zeros(n, dtype=np.uint64)
        assert self.lib.oracle_get_rng(self.h, _u64ptr(out), n) == 0
        return out


def xoroshiro_stream(s0: int, s1: int, n: int, model: str = "decay") -> np.ndarray:
    lib = _lib(model)
    s = np.array([s0, s1], dtype=np.uint64)
    out = np.zeros(n, dtype=np.uint64)
    lib.oracle_xoroshiro_next(_u64ptr(s), _u64ptr(out), n)
    return out


def randn_stream(s0: int, s1: int, n: int, precision: str = "double", model: str = "decay") -> np.ndarray:
    lib = _lib(model)
    s = np.array([s0, s1], dtype=np.uint64)
    out = np.zeros(n, dtype=np.float64)
    lib.oracle_randn(1 if precision == "single" else 2, _u64ptr(s), _dptr(out), n)
    return out


def cm_eval(fn: str, x, y=None, precision: str = "double", model: str = "decay") -> np.ndarray:
    lib = _lib(model)
    x = np.ascontiguousarray(x, dtype=np.float64)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(x)
    lib.oracle_cm_eval({"exp": 0, "log": 1, "pow": 2, "pown": 3, "root3": 4, "root5": 5, "min_step": 6, "root2": 7}[fn],
                       1 if precision == "single" else 2,
                       _dptr(x), None if yy is None else _dptr(yy), _dptr(out), x.size)
    return out
