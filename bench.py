#!/usr/bin/env python3
"""Benchmark of the ensemble-integration path on one MI355X.

Workload (one "step" = one pass of the hot path over the whole batch):
  lactotroph model, 256 x 256 grid over gK in [0, 10] x gBK in [0, 2]
  (65,536 members), x0 = (-60, 0, 0, 0.1), dopri5 rtol 1e-6 atol 1e-8,
  dt0 0.1, dtmax 1.0; transient over 5e3 ms, then the thresh2 observer
  (x_up 0.3, x_down 0.2) over 2e4 ms: warm-up pass + features pass.
Metric: member-steps/s (accepted steps of all members, counted on the
device, divided by the wall time of the K timed steps).  The headline runs in
FP64 (the reference's double-precision semantics, bit-exact against the CPU
oracle); the fp32-compensated mode is reported beside it with its feature
error against FP64.

Also reported: the roofline of the dominant kernel (features pass) from
HIP-event kernel times and the algorithmic FLOPs of clode_amd/perfmodel.py,
and a CPU baseline: the CPU oracle (the reference-semantics restatement,
not the reference binary, which cannot run here) timed on a member sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(REPO, "clode_amd", "models")


def lactotroph_sim(precision, n_side=256):
    import clode_amd as clode

    sim = clode.FeatureSimulator(
        src_file=os.path.join(MODELS, "lactotroph.cl"),
        variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
        parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0},
        aux=["ica"],
        observer=clode.Observer.threshold_2,
        stepper=clode.Stepper.dormand_prince,
        single_precision=precision,
        dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6, max_steps=100000000,
        t_span=(0.0, 5000.0),
        observer_x_up_thresh=0.3, observer_x_down_thresh=0.2,
    )
    gk = np.linspace(0.0, 10.0, n_side)
    gbk = np.linspace(0.0, 2.0, n_side)
    GK, GBK = np.meshgrid(gk, gbk, indexing="ij")  # member = j*n_side + k
    sim.set_ensemble(parameters={"gk": GK, "gbk": GBK})
    return sim


def one_step(sim):
    """transient 5e3, then thresh2 warm-up + features over 2e4 (x0 <- xf)."""
    sim.transient(t_span=(0.0, 5000.0))
    k_tr = sim.last_kernel_ms()
    sim.set_tspan((0.0, 20000.0))
    sim.initialize_observer()
    k_init = sim.last_kernel_ms()
    sim.features(initialize_observer=False, fetch_results=False)
    k_feat = sim.last_kernel_ms()
    return k_tr, k_init, k_feat


def run_lactotroph(precision, steps, warmup):
    sim = lactotroph_sim(precision)
    sim.enable_step_counter(True)
    for _ in range(warmup):
        one_step(sim)
    sim.reset_step_counts()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kernel_ms.append(one_step(sim))
    t1 = time.perf_counter()
    member_steps = int(sim.get_step_counts().sum())
    wall = t1 - t0
    # instrumented extra step: per-kernel member-steps for the roofline
    counts = []
    sim.reset_step_counts()
    sim.transient(t_span=(0.0, 5000.0))
    counts.append(int(sim.get_step_counts().sum()))
    k_tr = sim.last_kernel_ms()
    sim.set_tspan((0.0, 20000.0))
    sim.initialize_observer()
    counts.append(int(sim.get_step_counts().sum()) - counts[0])
    k_init = sim.last_kernel_ms()
    sim.features(initialize_observer=False, fetch_results=False)
    counts.append(int(sim.get_step_counts().sum()) - counts[0] - counts[1])
    k_feat = sim.last_kernel_ms()
    out = sim.get_observer_results()
    return {
        "member_steps": member_steps,
        "wall_s": wall,
        "member_steps_per_s": member_steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": {"transient": k_tr, "warmup": k_init, "features": k_feat},
        "kernel_member_steps": {"transient": counts[0], "warmup": counts[1], "features": counts[2]},
        "features": out.to_ndarray(),
        "feature_names": out.get_feature_names(),
    }


def feature_error(f32c, f64, names):
    """Relative error of fp32c features against FP64, for the period/peaks/V features."""
    res = {}
    for key in ("mean period", "mean peaks", "max v", "min v", "mean v", "event count"):
        j = names.index(key)
        a, b = f32c[:, j], f64[:, j]
        scale = np.maximum(np.abs(b), 1e-30)
        rel = np.abs(a - b) / scale
        valid = np.abs(b) > 0
        res[key] = {
            "median_rel": float(np.median(rel[valid])) if valid.any() else 0.0,
            "p99_rel": float(np.percentile(rel[valid], 99)) if valid.any() else 0.0,
            "frac_le_1e-4": float(np.mean(rel[valid] <= 1e-4)) if valid.any() else 1.0,
            "frac_gt_1pct": float(np.mean(rel[valid] > 1e-2)) if valid.any() else 0.0,
        }
    return res


def cpu_baseline(n_members=1024):
    """CPU oracle on every (65536/n)-th member, same sequence, all host threads."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle import OracleEnsemble

    threads = min(os.cpu_count() or 1, 16)
    gk = np.linspace(0.0, 10.0, 256)
    gbk = np.linspace(0.0, 2.0, 256)
    idx = np.arange(0, 65536, 65536 // n_members)
    pars = np.vstack([1.5 * np.ones(n_members), 3.0 * np.ones(n_members), gbk[idx % 256], gk[idx // 256]])
    x0 = np.tile(np.array([[-60.0], [0.0], [0.0], [0.1]]), (1, n_members))
    o = OracleEnsemble("lactotroph", 4, 4, 1, 0, stepper="dopri5", observer="thresh2", precision="double",
                       threads=threads)
    o.set_solver_params(0.1, 1.0, 1e-8, 1e-6, 100000000, 1000000, 1)
    o.set_observer_params(x_up_threshold=0.3, x_down_threshold=0.2)
    o.set_problem_data(x0, pars)
    t0 = time.perf_counter()
    o.set_tspan((0.0, 5000.0))
    o.transient()
    steps = o.last_steps
    o.shift_x0()
    o.set_tspan((0.0, 20000.0))
    o.features(1)
    steps += o.last_steps + o.last_warm_steps
    wall = time.perf_counter() - t0
    return {"value": steps / wall, "unit": "member-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_members} of 65536 members (every {65536 // n_members}th), same transient+features "
                      f"sequence, FP64, {steps} member-steps in {wall:.1f} s; CPU: {_cpu_model()}"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def lorenz_ratio(n=1 << 20, t_end=500.0):
    """Lorenz, 2^20 members, dopri5 + localmax: fp32c vs FP64 kernel time."""
    import clode_amd as clode

    rng = np.random.default_rng(20240601)
    x0 = rng.normal(1.0, 0.1, size=(n, 3))
    rho = np.linspace(0.0, 250.0, n)
    res = {}
    for prec in ("double", "fp32c"):
        sim = clode.FeatureSimulator(
            src_file=os.path.join(MODELS, "lorenz.cl"), variables={"x": 1.0, "y": 1.0, "z": 1.0},
            parameters={"sigma": 10.0, "rho": 28.0, "beta": 8.0 / 3.0}, observer=clode.Observer.local_max,
            stepper=clode.Stepper.dormand_prince, single_precision=prec, dt=0.01, dtmax=1.0, abstol=1e-8,
            reltol=1e-6, max_steps=100000000, t_span=(0.0, t_end))
        sim.set_ensemble(variables=x0, parameters={"rho": rho})
        sim.enable_step_counter(True)
        t0 = time.perf_counter()
        sim.features(fetch_results=False)
        wall = time.perf_counter() - t0
        steps = int(sim.get_step_counts().sum())
        res[prec] = {"kernel_ms": sim.last_kernel_ms(), "wall_s": wall, "member_steps": steps,
                     "member_steps_per_s": steps / wall}
    res["fp32c_speedup_kernel"] = res["double"]["kernel_ms"] / res["fp32c"]["kernel_ms"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-members", type=int, default=1024)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB device copy kernel (FETCH_SIZE/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from clode_amd.perfmodel import PEAK_TFLOPS, flops_per_step

    r64 = run_lactotroph("double", args.steps, args.warmup)
    print(f"# fp64: {r64['member_steps_per_s']:.4e} member-steps/s, {r64['ms_per_step']:.1f} ms/step, "
          f"kernels {r64['kernel_ms']}", file=sys.stderr, flush=True)
    fl = flops_per_step("lactotroph", "dopri5", "thresh2", 4, 1, "double")
    feat_steps = r64["kernel_member_steps"]["features"]
    feat_s = r64["kernel_ms"]["features"] / 1e3
    achieved = feat_steps * fl["flops"] / feat_s / 1e12
    roofline = {"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["f64"], "traffic": None,
                "kernel": "clode_features (thresh2, dopri5, FP64)",
                "flops_per_member_step": fl["flops"], "transcendentals_per_member_step": fl["transcendentals"],
                "member_steps_per_launch": feat_steps, "launch_ms": r64["kernel_ms"]["features"]}

    extra = {}
    if not args.no_extra:
        r32 = run_lactotroph("fp32c", args.steps, args.warmup)
        print(f"# fp32c: {r32['member_steps_per_s']:.4e} member-steps/s, {r32['ms_per_step']:.1f} ms/step",
              file=sys.stderr, flush=True)
        fl32 = flops_per_step("lactotroph", "dopri5", "thresh2", 4, 1, "fp32c")
        a32 = r32["kernel_member_steps"]["features"] * fl32["flops"] / (r32["kernel_ms"]["features"] / 1e3) / 1e12
        extra["fp32c"] = {
            "member_steps_per_s": r32["member_steps_per_s"], "ms_per_step": r32["ms_per_step"],
            "kernel_ms": r32["kernel_ms"], "speedup_vs_fp64_wall": r64["ms_per_step"] / r32["ms_per_step"],
            "features_kernel_tflops": a32, "features_kernel_frac_of_fp32_peak": a32 / PEAK_TFLOPS["f32"],
            "feature_error_vs_fp64": feature_error(r32["features"], r64["features"], r64["feature_names"]),
        }
        try:
            extra["lorenz_2p20_dopri5_localmax"] = lorenz_ratio()
        except Exception as e:  # extra figure only
            extra["lorenz_2p20_dopri5_localmax"] = {"error": str(e)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_members)
        print(f"# cpu baseline: {cpu['value']:.4e} member-steps/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    line = {
        "metric": "member-steps/s (lactotroph 256x256 feature map, dopri5, thresh2)",
        "value": r64["member_steps_per_s"],
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "lactotroph 256x256 gK[0,10] x gBK[0,2], dopri5 rtol 1e-6 atol 1e-8, "
                               "transient 5e3 ms + thresh2 warm-up and features 2e4 ms",
                   "members": 65536, "member_steps_per_step": r64["member_steps"] // max(args.steps, 1),
                   "kernel_ms": r64["kernel_ms"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
