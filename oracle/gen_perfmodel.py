"""Generates the operation table of clode_amd/perfmodel.py.

TEST / MEASUREMENT INFRASTRUCTURE (run by hand after a change to a stepper,
an observer or a model; tests/test_perfmodel.py checks that the committed
table is what this script produces).  Each benchmark workload's call
sequence runs on a deterministic sample of its members in the counting build
of the CPU oracle (oracle/flop_count.hpp: the oracle's steppers, controller,
observers, drivers and the model's right-hand side instantiated with an
operation-counting real type, FP64 semantics).  The counts are exact integers
for that sample; per-step figures divide them by the sample's steps and
trials.

Usage: python oracle/gen_perfmodel.py   (rewrites clode_amd/perfmodel.py)
"""
from __future__ import annotations

import os
import pprint
import sys

import numpy as np

ORACLE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ORACLE)
sys.path.insert(0, os.path.join(REPO, "tests"))
from oracle import OracleEnsemble  # noqa: E402

N2 = 1 << 20


def _lactotroph_grid(step):
    idx = np.arange(0, 65536, step)
    gk, gbk = np.linspace(0.0, 10.0, 256), np.linspace(0.0, 2.0, 256)
    pars = np.vstack([1.5 * np.ones(idx.size), 3.0 * np.ones(idx.size), gbk[idx % 256], gk[idx // 256]])
    return np.tile([[-60.0], [0.0], [0.0], [0.1]], (1, idx.size)), pars, None


def _vdp_c1(step):
    i = np.arange(0, 4096, step)
    mu = 0.1 + 4.9 * i / 4095
    return np.vstack([2.0 * np.ones(i.size), np.zeros(i.size)]), mu[None, :], None


def _noise(n):
    return (np.tile([[-60.0], [0.0], [0.0], [0.1]], (1, n)),
            np.vstack([1.5 * np.ones(n), 3.0 * np.ones(n), np.ones(n), 2.0 * np.ones(n)]), None)


def _lorenz(step, first=N2):
    idx = np.arange(0, first, step)
    x0 = np.random.default_rng(20240601).normal(1.0, 0.1, size=(N2, 3))[idx].T
    rho = 250.0 * idx / (N2 - 1)
    return x0, np.vstack([10.0 * np.ones(idx.size), rho, np.full(idx.size, 8.0 / 3.0)]), None


def _mixed(system, n, seed=7):
    """One C4 system: n members with spans 500 exp(z), z ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    span = np.vstack([np.zeros(n), 500.0 * np.exp(rng.standard_normal(n))])
    if system == "vdp":
        return np.vstack([2.0 * np.ones(n), np.zeros(n)]), rng.uniform(0.1, 5.0, n)[None, :], span
    if system == "lorenz":
        return (rng.normal(1.0, 0.1, (3, n)), np.vstack([10.0 * np.ones(n), rng.uniform(0, 50, n),
                                                          np.full(n, 8.0 / 3.0)]), span)
    if system == "ck":
        return (np.tile([[-50.0], [0.01], [0.12]], (1, n)),
                np.vstack([rng.uniform(550, 1050, n), 750.0 * np.ones(n), rng.uniform(0.095, 0.155, n)]), span)
    return np.zeros((1, n)), np.vstack([np.ones(n), np.zeros(n), 0.5 * np.ones(n)]), span


ADAPT = (0.1, 1.0, 1e-8, 1e-6, 10 ** 8, 10, 1)
LORENZ = (0.01, 1.0, 1e-8, 1e-6, 10 ** 8, 10, 1)
THRESH = dict(x_up_threshold=0.3, x_down_threshold=0.2)

# name -> model, (nVar, nPar, nAux, nNoise), stepper, observer, solver params,
# observer params, sample (x0, pars, per-member spans), sequence, sample text.
# Sequences: ("T", t0, t1) transient, ("S",) x0 <- xf, ("F", t0, t1, reinit)
# features (reinit 1: warm-up + features), ("J", t0, t1) trajectory.
WORKLOADS = {
    "lactotroph_256x256_dopri5_thresh2": dict(
        model="lactotroph", dims=(4, 4, 1, 0), stepper="dopri5", observer="thresh2", sp=ADAPT, obs=THRESH,
        sample=lambda: _lactotroph_grid(1024), seq=[("T", 0.0, 5000.0), ("S",), ("F", 0.0, 20000.0, 1)],
        text="64 of the 65,536 grid members (every 1024th), transient 5e3 ms, warm-up and features 2e4 ms"),
    "vdp_4096_rk4_thresh2": dict(
        model="van_der_pol", dims=(2, 1, 0, 0), stepper="rk4", observer="thresh2",
        sp=(0.01, 0.01, 1e-6, 1e-3, 10 ** 8, 10, 1), obs=THRESH,
        sample=lambda: _vdp_c1(256), seq=[("T", 0.0, 1000.0), ("S",), ("F", 0.0, 1000.0, 1)],
        text="16 of the 4,096 members (every 256th), transient 1e3, warm-up and features 1e3"),
    "lactotroph_noise_seuler_thresh2": dict(
        model="lactotroph_noise", dims=(4, 4, 1, 1), stepper="seuler", observer="thresh2",
        sp=(0.05, 0.05, 1e-6, 1e-3, 10 ** 8, 10, 1), obs=THRESH, seed=2024,
        sample=lambda: _noise(4), seq=[("T", 0.0, 5000.0), ("S",), ("F", 0.0, 20000.0, 1)],
        text="4 identical members (sigma 2), seed 2024, transient 5e3 ms, warm-up and features 2e4 ms"),
    "lorenz_2p20_dopri5_localmax": dict(
        model="lorenz", dims=(3, 3, 0, 0), stepper="dopri5", observer="localmax", sp=LORENZ, obs={},
        sample=lambda: _lorenz(1 << 16), seq=[("F", 0.0, 500.0, 1)],
        text="16 of the 2^20 members (every 65,536th: rho 0 to 234), features over (0, 500)"),
    "lorenz_2p20_bs23_localmax": dict(
        model="lorenz", dims=(3, 3, 0, 0), stepper="bs23", observer="localmax", sp=LORENZ, obs={},
        sample=lambda: _lorenz(1 << 16), seq=[("F", 0.0, 500.0, 1)],
        text="16 of the 2^20 members (every 65,536th: rho 0 to 234), features over (0, 500)"),
    "lorenz_65536x1024_dopri5_trajectory": dict(
        model="lorenz", dims=(3, 3, 0, 0), stepper="dopri5", observer="", sp=(0.01, 1.0, 1e-8, 1e-6, 10 ** 8, 1024, 1),
        obs=None, kind="trajectory", sample=lambda: _lorenz(1024, first=65536), seq=[("J", 0.0, 500.0)],
        text="64 of the first 65,536 members (every 1024th), 1024 samples, nout 1"),
    "mixed_vdp_dopri5_thresh2": dict(
        model="van_der_pol", dims=(2, 1, 0, 0), stepper="dopri5", observer="thresh2",
        sp=(0.01, 1.0, 1e-8, 1e-6, 10 ** 9, 10, 1), obs=THRESH, seed=100,
        sample=lambda: _mixed("vdp", 16), seq=[("T", 0.0, 1.0), ("S",), ("F", 0.0, 1.0, -1), ("S",)],
        text="16 members, spans 500 exp(z), one round (transient, features)"),
    "mixed_lorenz_dopri5_localmax": dict(
        model="lorenz", dims=(3, 3, 0, 0), stepper="dopri5", observer="localmax",
        sp=(0.01, 1.0, 1e-8, 1e-6, 10 ** 9, 10, 1), obs={}, seed=101,
        sample=lambda: _mixed("lorenz", 16), seq=[("T", 0.0, 1.0), ("S",), ("F", 0.0, 1.0, -1), ("S",)],
        text="16 members, rho in [0, 50], spans 500 exp(z), one round (transient, features)"),
    "mixed_chay_keizer_dopri5_thresh2": dict(
        model="chay_keizer", dims=(3, 3, 0, 0), stepper="dopri5", observer="thresh2",
        sp=(0.01, 1.0, 1e-8, 1e-6, 10 ** 9, 10, 1), obs=THRESH, seed=102,
        sample=lambda: _mixed("ck", 16), seq=[("T", 0.0, 1.0), ("S",), ("F", 0.0, 1.0, -1), ("S",)],
        text="16 members, spans 500 exp(z), one round (transient, features)"),
    "mixed_ou_seuler_basicall": dict(
        model="ornstein_uhlenbeck", dims=(1, 3, 0, 1), stepper="seuler", observer="basicall",
        sp=(0.05, 0.05, 1e-8, 1e-6, 10 ** 9, 10, 1), obs={}, seed=103,
        sample=lambda: _mixed("ou", 16), seq=[("T", 0.0, 1.0), ("S",), ("F", 0.0, 1.0, -1), ("S",)],
        text="16 members, spans 500 exp(z), one round (transient, features)"),
}

FLOP_KEYS = ("add", "mul", "div", "fma")
TRANS_KEYS = ("exp", "log", "pow", "pown", "sqrt", "other")


def flops(c: dict) -> int:
    return c["add"] + c["mul"] + c["div"] + 2 * c["fma"]


def count_workload(name: str) -> dict:
    """Exact operation counts of one workload's sample, per phase."""
    w = WORKLOADS[name]
    nv, npar, na, nn = w["dims"]
    o = OracleEnsemble(w["model"], nv, npar, na, nn, stepper=w["stepper"], observer=w["observer"],
                       precision="count", kind=w.get("kind", "features" if w["observer"] else "base"))
    o.set_solver_params(*w["sp"])
    if w["observer"]:
        o.set_observer_params(**w["obs"])
    x0, pars, span = w["sample"]()
    o.set_problem_data(x0, pars)
    if span is not None:
        o.set_member_tspan(span)
    if "seed" in w:
        o.seed_rng(w["seed"])
    for op in w["seq"]:
        if op[0] == "S":
            o.shift_x0()
            continue
        o.set_tspan((op[1], op[2]))
        if op[0] == "T":
            o.transient()
        elif op[0] == "F":
            o.features(op[3])
        else:
            o.trajectory()
    out = {}
    for phase, d in o.counts().items():
        t = d["trials"]
        steps = t["accepted"] + t["failed"]
        if steps == 0:
            continue
        trial_flops = {b: flops(d[b]) for b in ("accepted", "rejected", "failed")}
        other = flops(d["all"]) - sum(trial_flops.values())
        out[phase] = {
            "steps": steps,
            "trials": t,
            "flops": dict(trial_flops, other=other),
            "ops": {k: d["all"][k] for k in FLOP_KEYS},
            "transcendentals": {k: d["all"][k] for k in TRANS_KEYS if d["all"][k]},
        }
    return out


def per_step(phase: dict) -> dict:
    t, f = phase["trials"], phase["flops"]
    acc = f["accepted"] / max(t["accepted"], 1)
    return {
        "accepted_trial": round(acc, 3),
        "rejected_trial": round(f["rejected"] / t["rejected"], 3) if t["rejected"] else round(acc, 3),
        "other": round(f["other"] / phase["steps"], 3),
        "step": round((f["accepted"] + f["failed"] + f["other"]) / phase["steps"], 3),
        "rejected_per_step": round(t["rejected"] / phase["steps"], 5),
        "transcendentals": {k: round(v / phase["steps"], 4) for k, v in phase["transcendentals"].items()},
    }


HEADER = '''"""Algorithmic operation counts of the integration path, for roofline accounting.

GENERATED by oracle/gen_perfmodel.py; do not edit by hand
(tests/test_perfmodel.py checks that this table is what the generator
produces from the current sources).

A member-step is one completed step of one ensemble member (the reference's
``step`` counter, clode/cpp/transient.cl:54-58): an accepted step, or a step
that failed at hmin.  A step trial is one stepper call: an accepted step plus
each rejected trial of the adaptive controller.  The device counts both per
member (clode_get_step_counts / clode_get_trial_counts).

The counts come from the CPU oracle's own steppers, controller, observers,
drivers and the models' right-hand sides, instantiated with an
operation-counting real type (oracle/flop_count.hpp), over a fixed sample of
each workload's members (``sample``):
  * +, -, *, / count 1 and fma counts 2 when at least one operand varies
    during the integration (state, time, step size, stage values, observer
    state, random numbers); operations on constants, tableau coefficients and
    solver, observer or model parameters only are folded or hoisted by a
    compiler and count 0; negation, fabs, fmin, fmax, comparisons and integer
    work count 0;
  * exp, log, pow, pown and sqrt are listed separately, per member-step, and
    are not counted as FLOPs.
Per phase (transient, warm-up, features, trajectory): exact totals for the
sample, and per-step figures: ``accepted_trial`` (the step set-up and the
accepted trial), ``rejected_trial``, ``other`` (observer update and events,
per-member set-up, finalisation, per member-step) and ``step`` (accepted
trial + other).  The counts follow the reference's FP64 semantics; the
float-compensated mode runs the same algorithm (its compensation and double
time are not algorithmic work).
"""
from __future__ import annotations

'''

FOOTER = '''

def phase_counts(workload: str, phase: str) -> dict:
    """The per-step figures of one workload's phase."""
    return TABLE[workload]["phases"][phase]["per_step"]


def launch_flops(workload: str, phase: str, member_steps: int, trials: int | None = None) -> float:
    """Algorithmic FLOPs of a launch that completed ``member_steps`` steps in
    ``trials`` step trials (both counted on the device): every step costs
    ``step``, every rejected trial ``rejected_trial``.  Without a trial
    count the sample's rejected trials per step stand in."""
    p = phase_counts(workload, phase)
    rejected = (trials - member_steps) if trials is not None else member_steps * p["rejected_per_step"]
    return member_steps * p["step"] + max(rejected, 0) * p["rejected_trial"]


# MI355X published peaks (dense)
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16_mfma": 2500.0}
PEAK_HBM_GBS = 8000.0
'''


def generate() -> dict:
    table = {}
    for name, w in WORKLOADS.items():
        phases = count_workload(name)
        for ph in phases.values():
            ph["per_step"] = per_step(ph)
        table[name] = {"model": w["model"], "stepper": w["stepper"], "observer": w["observer"],
                       "sample": w["text"], "phases": phases}
        print(name, {k: v["per_step"]["step"] for k, v in phases.items()}, flush=True)
    return table


def render(table: dict) -> str:
    return HEADER + "TABLE = " + pprint.pformat(table, width=116, sort_dicts=False) + "\n" + FOOTER


if __name__ == "__main__":
    text = render(generate())
    with open(os.path.join(REPO, "clode_amd", "perfmodel.py"), "w") as f:
        f.write(text)
    print("wrote clode_amd/perfmodel.py")
