"""Drive the GPU simulator (pybind11 module over the C ABI) and the CPU oracle
with the same call sequence, and compare their outputs."""
from __future__ import annotations

import os

import numpy as np

from oracle import OracleEnsemble

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(REPO, "clode_amd", "models")

# model name -> (vars, pars, aux, noise)
MODEL_INFO = {
    "van_der_pol": (["x", "y"], ["mu"], [], 0),
    "lorenz": (["x", "y", "z"], ["sigma", "rho", "beta"], [], 0),
    "lactotroph": (["v", "n", "f", "c"], ["gcal", "gsk", "gbk", "gk"], ["ica"], 0),
    "lactotroph_noise": (["v", "n", "f", "c"], ["gcal", "gsk", "gbk", "sigma"], ["ica"], 1),
    "chay_keizer": (["v", "n", "c"], ["gca", "gkca", "kpmca"], [], 0),
    "ornstein_uhlenbeck": (["x"], ["theta", "mu", "sigma"], [], 1),
    "sine": (["x"], ["omega"], ["xp1", "pos", "neg"], 0),
    "linear_drag": (["y1", "y2"], ["m", "w", "k", "H"], ["g1"], 0),
    "harmonic": (["x", "y"], ["omega"], [], 0),
    "decay": (["x"], ["lambda"], [], 0),
    "two_tone": (["u1", "v1", "u2", "v2", "s"], ["b"], [], 0),
}

OBS_DEFAULTS = dict(e_var_ix=0, f_var_ix=0, max_event_count=100, max_event_timestamps=0, min_amp=0.0, min_imi=0.0,
                    nhood_radius=0.05, x_up_threshold=0.3, x_down_threshold=0.2, dx_up_threshold=0.0,
                    dx_down_threshold=0.0, eps_dx=0.0)


def wrapper():
    from clode_amd.cpp import clode_cpp_wrapper as w

    return w


class Pair:
    """A GPU simulator and its oracle twin."""

    def __init__(self, model, stepper="rk4", observer="", precision="double", kind="features", obs=None,
                 sp=(0.1, 1.0, 1e-6, 1e-3, 1000000, 1000000, 1), oracle=True):
        w = wrapper()
        vars_, pars, aux, noise = MODEL_INFO[model]
        self.model, self.kind, self.precision = model, kind, precision
        self.nv, self.np, self.na = len(vars_), len(pars), len(aux)
        self.obs = dict(OBS_DEFAULTS)
        if obs:
            self.obs.update(obs)
        pi = w.ProblemInfo(os.path.join(MODELS, model + ".cl"), vars_, pars, aux, noise)
        dev = w.OpenCLResource()
        root = os.path.join(REPO, "clode_amd", "_jitcache")
        if kind == "features":
            self.op = w.ObserverParams(**self.obs)
            self.gpu = w.FeatureSimulatorBase(pi, stepper, observer, self.op, precision, dev, root)
        elif kind == "trajectory":
            self.gpu = w.TrajectorySimulatorBase(pi, stepper, precision, dev, root)
        else:
            self.gpu = w.SimulatorBase(pi, stepper, precision, dev, root)
        self.gpu.build_cl()
        self.sp = sp
        self.gpu.set_solver_params(w.SolverParams(*sp))
        self.orc = None
        if oracle:
            self.orc = OracleEnsemble(model, self.nv, self.np, self.na, noise, stepper=stepper,
                                      observer=observer if kind == "features" else "",
                                      precision=precision, n_store=self.obs["max_event_timestamps"], kind=kind)
            self.orc.set_solver_params(*sp)
            if kind == "features":
                self.orc.set_observer_params(**self.obs)

    # ---- identical calls on both sides
    def set_problem_data(self, x0, pars):
        x0 = np.asarray(x0, dtype=np.float64)
        pars = np.asarray(pars, dtype=np.float64)
        self.n = x0.shape[1]
        self.gpu.set_problem_data(list(x0.ravel()), list(pars.ravel()))
        if self.orc:
            self.orc.set_problem_data(x0, pars)

    def set_tspan(self, t0, t1):
        self.gpu.set_tspan([t0, t1])
        if self.orc:
            self.orc.set_tspan((t0, t1))

    def set_solver_params(self, sp):
        w = wrapper()
        self.sp = sp
        self.gpu.set_solver_params(w.SolverParams(*sp))
        if self.orc:
            self.orc.set_solver_params(*sp)

    def set_member_tspan(self, tm):
        """Per-member spans [2][n] (start row, end row) on both sides."""
        tm = np.asarray(tm, dtype=np.float64)
        self.gpu.set_member_tspan(list(tm.ravel()))
        if self.orc:
            self.orc.set_member_tspan(tm)

    def seed(self, s):
        self.gpu.seed_rng(s)
        if self.orc:
            self.orc.seed_rng(s)

    def transient(self):
        self.gpu.transient()
        if self.orc:
            self.orc.transient()

    def initialize_observer(self):
        self.gpu.initialize_observer()
        if self.orc:
            self.orc.initialize_observer()

    def features(self, reinit=None):
        if reinit is None:
            self.gpu.features()
        else:
            self.gpu.features(bool(reinit))
        if self.orc:
            self.orc.features(-1 if reinit is None else int(bool(reinit)))

    def trajectory(self):
        self.gpu.trajectory()
        if self.orc:
            self.orc.trajectory()

    def shift_x0(self):
        self.gpu.shift_x0()
        if self.orc:
            self.orc.shift_x0()

    # ---- results
    def gpu_get(self, name):
        g = self.gpu
        return np.array({"xf": g.get_xf, "x0": g.get_x0, "dt": g.get_dt, "tf": g.get_tf,
                         "F": getattr(g, "get_f", None), "t": getattr(g, "get_t", None),
                         "x": getattr(g, "get_x", None), "dx": getattr(g, "get_dx", None),
                         "aux": getattr(g, "get_aux", None), "nStored": getattr(g, "get_n_stored", None)}[name](),
                        dtype=np.float64)

    def compare(self, names=("xf", "dt", "tf")):
        """Return {name: (n_bitexact, n_total, max_abs)}."""
        out = {}
        for name in names:
            a = self.gpu_get(name)
            b = self.orc.get(name)
            assert a.shape == b.shape, (name, a.shape, b.shape)
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            d = np.abs(a - b)
            out[name] = (int(same.sum()), a.size, float(np.nanmax(d)) if a.size else 0.0)
        return out

    def assert_bitexact(self, names=("xf", "dt", "tf")):
        res = self.compare(names)
        for name, (same, total, maxabs) in res.items():
            assert same == total, f"{name}: {same}/{total} bit-exact, max |diff| {maxabs:g}"
        return res

    def assert_trajectory_bitexact(self):
        """Compare the stored samples only: slot s of member i is defined for
        s < min(nStored[i], max_store); the rest of each buffer is whatever
        the device memory held (uninitialised in the reference as well)."""
        self.assert_bitexact(("nStored", "xf", "dt", "tf"))
        ns = self.gpu_get("nStored").astype(np.int64)
        store = self.sp[5]
        valid = np.arange(store)[:, None] < np.minimum(ns, store)[None, :]  # [slot][member]
        for name, width in (("t", 1), ("x", self.nv), ("dx", self.nv), ("aux", self.na)):
            if width == 0:
                continue
            a = self.gpu_get(name).reshape(store, width, self.n)
            b = self.orc.get(name).reshape(store, width, self.n)
            m = np.broadcast_to(valid[:, None, :], a.shape)
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            bad = int((~same & m).sum())
            assert bad == 0, f"{name}: {bad}/{int(m.sum())} stored samples differ"
        return int(valid.sum())
