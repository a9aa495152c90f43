"""Drive the MATLAB gateways (clode_amd/matlab/clode_mex.cpp) in-process
through the stand-in MEX runtime of tests/mex_stub: build MATLAB values,
call the gateway with an action string, read the outputs."""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "mex_stub"))
import build_fake_mex  # noqa: E402

_libs = {}


class MexError(RuntimeError):
    pass


class Gateway:
    """One gateway: kind 0 clODEmex, 1 clODEfeaturesmex, 2 clODEtrajectorymex."""

    def __init__(self, kind: int):
        if kind not in _libs:
            path = build_fake_mex.lib_path(kind)
            if not path.exists():
                build_fake_mex.build(kind)
            lib = ctypes.CDLL(str(path))
            vp, sz, dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)
            lib.fake_double.restype = vp
            lib.fake_double.argtypes = [dp, sz, sz]
            lib.fake_string.restype = vp
            lib.fake_string.argtypes = [ctypes.c_char_p]
            lib.fake_cell.restype = vp
            lib.fake_cell.argtypes = [sz]
            lib.fake_cell_set.argtypes = [vp, sz, vp]
            lib.fake_struct.restype = vp
            lib.fake_struct_set.argtypes = [vp, ctypes.c_char_p, vp]
            lib.fake_call.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
            lib.fake_last_error.restype = ctypes.c_char_p
            lib.fake_numel.restype = sz
            lib.fake_numel.argtypes = [vp]
            lib.fake_data.restype = dp
            lib.fake_data.argtypes = [vp]
            lib.fake_is_cell.argtypes = [vp]
            lib.fake_cell_get.restype = vp
            lib.fake_cell_get.argtypes = [vp, sz]
            lib.fake_str.restype = ctypes.c_char_p
            lib.fake_str.argtypes = [vp]
            _libs[kind] = lib
        self.lib = _libs[kind]

    # ---- MATLAB values
    def value(self, v):
        lib = self.lib
        if isinstance(v, str):
            return lib.fake_string(v.encode())
        if isinstance(v, dict):
            s = lib.fake_struct()
            for k, x in v.items():
                lib.fake_struct_set(s, k.encode(), self.value(x))
            return s
        if isinstance(v, (list, tuple)) and v and isinstance(v[0], str):
            c = lib.fake_cell(len(v))
            for i, x in enumerate(v):
                lib.fake_cell_set(c, i, self.value(x))
            return c
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))
        return lib.fake_double(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, a.size)

    def result(self, p):
        lib = self.lib
        if not p:
            return None
        if lib.fake_is_cell(p):
            return [lib.fake_str(lib.fake_cell_get(p, i)).decode() for i in range(lib.fake_numel(p))]
        n = lib.fake_numel(p)
        return np.ctypeslib.as_array(lib.fake_data(p), shape=(n,)).copy() if n else np.zeros(0)

    def __call__(self, action: str, *args, nout: int = 1):
        vals = [self.value(action)] + [self.value(a) for a in args]
        prhs = (ctypes.c_void_p * len(vals))(*vals)
        plhs = (ctypes.c_void_p * max(nout, 1))()
        if self.lib.fake_call(nout, plhs, len(vals), prhs) != 0:
            raise MexError(self.lib.fake_last_error().decode())
        return self.result(plhs[0])

    def lock_count(self) -> int:
        return self.lib.fake_lock_count()


def problem_struct(path, vars_, pars, aux, noise):
    s = {"clRHSfilename": path, "nVar": len(vars_), "nPar": len(pars), "nAux": len(aux), "nWiener": noise,
         "varNames": vars_, "parNames": pars}
    if aux:
        s["auxNames"] = aux
    return s


def solver_struct(dt, dtmax, abstol, reltol, max_steps, max_store, nout):
    return {"dt": dt, "dtmax": dtmax, "abstol": abstol, "reltol": reltol, "max_steps": max_steps,
            "max_store": max_store, "nout": nout}
