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

Extra workloads (--extras, default all): fp32c (the headline in fp32c), seq
(the step as one launch), replay (the step with the features pass replaying
the warm-up's recorded steps from HBM), c1
(Van der Pol 4096, rk4), noise (the lactotroph Euler-Maruyama variant),
c3 (Lorenz 2^20, and the 65,536 x 1024-sample trajectory), c4 (mixed 2^20
batched sweep), c5 (surrogate-screened 2048 x 2048 map).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra] [--extras LIST]
"""
from __future__ import annotations

import argparse
import inspect
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(REPO, "clode_amd", "models")


def lactotroph_sim(precision, n_side=256, rtol=1e-6, atol=1e-8):
    import clode_amd as clode

    sim = clode.FeatureSimulator(
        src_file=os.path.join(MODELS, "lactotroph.cl"),
        variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
        parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0},
        aux=["ica"],
        observer=clode.Observer.threshold_2,
        stepper=clode.Stepper.dormand_prince,
        single_precision=precision,
        dt=0.1, dtmax=1.0, abstol=atol, reltol=rtol, max_steps=100000000,
        t_span=(0.0, 5000.0),
        observer_x_up_thresh=0.3, observer_x_down_thresh=0.2,
    )
    gk = np.linspace(0.0, 10.0, n_side)
    gbk = np.linspace(0.0, 2.0, n_side)
    GK, GBK = np.meshgrid(gk, gbk, indexing="ij")  # member = j*n_side + k
    sim.set_ensemble(parameters={"gk": GK, "gbk": GBK})
    return sim


def one_step(sim):
    """transient 5e3, then thresh2 warm-up + features over 2e4 (x0 <- xf),
    and the feature matrix read back to the host, as the reference's
    features() does (clode/features.py:467-507)."""
    sim.transient(t_span=(0.0, 5000.0))
    k_tr = sim.last_kernel_ms()
    sim.set_tspan((0.0, 20000.0))
    sim.initialize_observer()
    k_init = sim.last_kernel_ms()
    sim.features(initialize_observer=False, fetch_results=True)
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
    # instrumented extra step: per-kernel member-steps and step trials (steps
    # plus rejected trials) for the roofline
    counts, trials = [], []

    def take():
        s, t = int(sim.get_step_counts().sum()), int(sim.get_trial_counts().sum())
        counts.append(s - sum(counts))
        trials.append(t - sum(trials))

    sim.reset_step_counts()
    x0_step = sim.get_initial_state().copy()
    sim.transient(t_span=(0.0, 5000.0))
    take()
    k_tr = sim.last_kernel_ms()
    sim.set_tspan((0.0, 20000.0))
    sim.initialize_observer()
    take()
    k_init = sim.last_kernel_ms()
    sim.features(initialize_observer=False, fetch_results=False)
    take()
    k_feat = sim.last_kernel_ms()
    out = sim.get_observer_results()
    return {
        "member_steps": member_steps,
        "wall_s": wall,
        "member_steps_per_s": member_steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": {"transient": k_tr, "warmup": k_init, "features": k_feat},
        "kernel_member_steps": {"transient": counts[0], "warmup": counts[1], "features": counts[2]},
        "kernel_trials": {"transient": trials[0], "warmup": trials[1], "features": trials[2]},
        "features": out.to_ndarray(),
        "feature_names": out.get_feature_names(),
        "x0_step": x0_step,
    }


def lactotroph_step_features(precision, x0, rtol=1e-6, atol=1e-8):
    """Features of one bench step (transient, warm-up, features) of a fresh
    simulator from the initial states x0: the fp32c error study runs FP64,
    fp32c and FP64 at rtol / 100 (the yardstick of the method's own
    truncation error) from the same states and the same first step size."""
    sim = lactotroph_sim(precision, rtol=rtol, atol=atol)
    gk, gbk = np.linspace(0.0, 10.0, 256), np.linspace(0.0, 2.0, 256)
    GK, GBK = np.meshgrid(gk, gbk, indexing="ij")
    sim.set_ensemble(variables=x0, parameters={"gk": GK, "gbk": GBK})
    one_step(sim)
    return sim.get_observer_results().to_ndarray()


def one_launch_sequence(precision, steps, r_three=None):
    """The same step as one_step() with the transient, x0 <- xf, warm-up and
    features of every member in one launch (clode_transient_features): wall
    time per step with the feature readback, the kernel time, and whether the
    features equal those of the three launches bit for bit."""
    sim = lactotroph_sim(precision)
    sim.transient_features((0.0, 5000.0), (0.0, 20000.0), fetch_results=False)  # program load, untimed
    sim = lactotroph_sim(precision)
    first = sim.transient_features((0.0, 5000.0), (0.0, 20000.0)).to_ndarray()
    kernel_ms = [sim.last_kernel_ms()]
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.transient_features((0.0, 5000.0), (0.0, 20000.0))
        kernel_ms.append(sim.last_kernel_ms())
    wall = time.perf_counter() - t0
    ref = lactotroph_sim(precision)
    one_step(ref)
    same = bool(np.array_equal(first, ref.get_observer_results().to_ndarray()))
    res = {"ms_per_step": 1e3 * wall / steps, "kernel_ms": float(np.median(kernel_ms[1:])),
           "identical_features_to_three_launches": same}
    if r_three is not None:
        three = sum(r_three["kernel_ms"].values())
        res["three_launch_kernel_ms"] = three
        res["kernel_speedup"] = three / res["kernel_ms"]
        res["wall_speedup"] = r_three["ms_per_step"] / res["ms_per_step"]
    return res


def replay_step(precision, steps, r_plain=None):
    """The same step with the features replay on: the thresh2 warm-up pass
    records every member's accepted steps in HBM (time, state, derivative,
    aux, step size as doubles: 88 B per member-step in FP64, 120 B in fp32c)
    and the features pass reads them back instead of integrating them again.
    Wall time per step with the feature readback, the kernel times, the
    record's HBM bytes and read rate, the member-steps actually integrated,
    and whether the features equal those of the integrating step bit for
    bit (both from the same initial states)."""
    rec = 8 * (2 + 2 * 4 + 1 + (4 if precision == "fp32c" else 0))
    budget = 65536 * rec * 24000  # every member's warm-up (at most ~21,600 steps here)
    ref = lactotroph_sim(precision)
    one_step(ref)
    want = ref.get_observer_results().to_ndarray()
    del ref
    sim = lactotroph_sim(precision)
    sim.set_replay_budget(budget)
    sim.enable_step_counter(True)
    one_step(sim)
    same = bool(np.array_equal(sim.get_observer_results().to_ndarray(), want))
    cap, replayed = sim.replay_info()
    sim.reset_step_counts()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kernel_ms.append(one_step(sim))
    wall = time.perf_counter() - t0
    integrated = int(sim.get_step_counts().sum())
    # one more step, instrumented: the warm-up's recorded steps
    sim.reset_step_counts()
    sim.transient(t_span=(0.0, 5000.0))
    tr = int(sim.get_step_counts().sum())
    sim.set_tspan((0.0, 20000.0))
    sim.initialize_observer()
    warm = int(sim.get_step_counts().sum()) - tr
    k_rec = sim.last_kernel_ms()
    sim.features(initialize_observer=False, fetch_results=False)
    k_rep = sim.last_kernel_ms()
    sim.set_replay_budget(0)
    k = np.median(np.array(kernel_ms), axis=0)
    res = {"ms_per_step": 1e3 * wall / steps,
           "kernel_ms": {"transient": float(k[0]), "warmup_record": float(k[1]), "features_replay": float(k[2])},
           "integrated_member_steps_per_s": integrated / wall,
           "records_per_member": cap, "replayed": replayed,
           "record_bytes": warm * rec, "record_bytes_per_member_step": rec,
           "replay_read_GBps": warm * rec / (k_rep / 1e3) / 1e9,
           "replay_read_frac_of_hbm_peak": warm * rec / (k_rep / 1e3) / 1e9 / 8000.0,
           "record_write_GBps_during_warmup": warm * rec / (k_rec / 1e3) / 1e9,
           "identical_features_to_integrating_step": same}
    if r_plain is not None:
        res["wall_speedup_vs_integrating_step"] = r_plain["ms_per_step"] / res["ms_per_step"]
        res["features_kernel_speedup"] = r_plain["kernel_ms"]["features"] / res["kernel_ms"]["features_replay"]
    return res


def kernel_tflops(r, phase, workload="lactotroph_256x256_dopri5_thresh2"):
    """Algorithmic TFLOP/s of one launch of the timed sequence: the
    generated operation table (clode_amd/perfmodel.py) priced with the
    launch's device-counted steps and trials, over its HIP-event time."""
    from clode_amd.perfmodel import launch_flops

    fl = launch_flops(workload, phase, r["kernel_member_steps"][phase], r["kernel_trials"][phase])
    return fl / (r["kernel_ms"][phase] / 1e3) / 1e12, fl


def feature_error(f32c, f64, names, fine=None):
    """Relative error of every fp32c feature against FP64 over the members:
    median, p99, the fraction within 1e-4 and the fraction beyond 1 %.  With
    `fine` (FP64 at rtol / 100 on the same members), also the same figures
    for FP64's own truncation error |fp64 - fine| and the fraction of members
    on which fp32c is within 3x of it (plus 1e-5 (1 + |fp64|))."""
    res = {}
    for key in names:
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
        if fine is not None:
            t = np.abs(b - fine[:, j])
            trel = t / scale
            res[key]["fp64_truncation_median_rel"] = float(np.median(trel[valid])) if valid.any() else 0.0
            res[key]["fp64_truncation_p99_rel"] = float(np.percentile(trel[valid], 99)) if valid.any() else 0.0
            res[key]["frac_within_3x_fp64_truncation"] = float(
                np.mean(np.abs(a - b) <= 3 * t + 1e-5 * (1 + np.abs(b))))
    return res


def job_cores():
    """The cores this process may use: its CPU affinity, capped by a cgroup
    CPU quota when one is set (a shared GPU host gives each job a share of
    its cores), with the host's core count beside them."""
    host = os.cpu_count() or 1
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else host
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    if quota:
        cores = max(1, min(cores, int(quota)))
    return cores, host, quota


def cpu_baseline(n_members=1024):
    """CPU oracle on every (65536/n)-th member, same sequence, every core
    the job may use."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle import OracleEnsemble

    threads, host, quota = job_cores()
    threads = min(threads, int(os.environ.get("CLODE_BENCH_CPU_THREADS", threads)))
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
            "host_cpus": host, "cpu_quota": quota,
            "sample": f"{n_members} of 65536 members (every {65536 // n_members}th), same transient+features "
                      f"sequence, FP64, {steps} member-steps in {wall:.1f} s on {threads} threads "
                      f"({wall * threads:.0f} core-seconds; every core of the job's share: {threads} of the "
                      f"host's {host}); CPU: {_cpu_model()}"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def vdp_c1(n=4096, cpu_members=256):
    """Van der Pol, 4096 members, mu in [0.1, 5], rk4 dt 0.01, FP64: transient
    over 1e3, x0 <- xf, thresh2 (warm-up + features) over 1e3.  GPU
    member-steps/s, and the CPU oracle on every (n / cpu_members)-th member."""
    import clode_amd as clode

    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "van_der_pol.cl"), variables={"x": 2.0, "y": 0.0},
                                 parameters={"mu": 1.0}, observer=clode.Observer.threshold_2,
                                 stepper=clode.Stepper.rk4, single_precision=False, dt=0.01, dtmax=0.01,
                                 max_steps=10 ** 8, t_span=(0.0, 1000.0), observer_x_up_thresh=0.3,
                                 observer_x_down_thresh=0.2)
    mu = 0.1 + 4.9 * np.arange(n) / (n - 1)
    sim.set_ensemble(parameters={"mu": mu})
    sim.enable_step_counter(True)
    sim.transient(t_span=(0.0, 1000.0))  # warm (compile, first launch)
    sim.set_ensemble(parameters={"mu": mu})
    sim.reset_step_counts()
    t0 = time.perf_counter()
    sim.transient(t_span=(0.0, 1000.0))
    out = sim.features(t_span=(0.0, 1000.0))
    wall = time.perf_counter() - t0
    steps = int(sim.get_step_counts().sum())
    res = {"members": n, "member_steps": steps, "wall_s": wall, "member_steps_per_s": steps / wall,
           "mean_period_mu1": float(np.interp(1.0, mu, np.ravel(out.get_var_max("period"))))}
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle import OracleEnsemble

    idx = np.arange(0, n, n // cpu_members)
    threads = job_cores()[0]
    o = OracleEnsemble("van_der_pol", 2, 1, 0, 0, stepper="rk4", observer="thresh2", precision="double",
                       threads=threads)
    o.set_solver_params(0.01, 0.01, 1e-6, 1e-3, 10 ** 8, 1000000, 1)
    o.set_observer_params(x_up_threshold=0.3, x_down_threshold=0.2)
    o.set_problem_data(np.vstack([2.0 * np.ones(idx.size), np.zeros(idx.size)]), mu[idx][None, :])
    t0 = time.perf_counter()
    o.set_tspan((0.0, 1000.0))
    o.transient()
    cs = o.last_steps
    o.shift_x0()
    o.features(1)
    cs += o.last_steps + o.last_warm_steps
    cw = time.perf_counter() - t0
    res["cpu_oracle"] = {"members": int(idx.size), "threads": threads, "member_steps_per_s": cs / cw}
    return res


def noise_c2(n=65536, seed=2024):
    """The lactotroph noise variant: 65,536 identical members (gcal 1.5, gsk 3,
    gbk 1, sigma 2 in lactotroph_noise.cl), Euler-Maruyama dt 0.05, seeded;
    transient over 5e3 ms, then thresh2 warm-up and features over 2e4 ms, in
    FP64 and fp32c.  Member-steps/s of the sequence and the spread of the
    mean period across the noise realisations."""
    import clode_amd as clode

    res = {}
    for prec in ("double", "fp32c"):
        sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph_noise.cl"),
                                     variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                     parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "sigma": 2.0}, aux=["ica"],
                                     num_noise=1, observer=clode.Observer.threshold_2,
                                     stepper=clode.Stepper.stochastic_euler, single_precision=prec, dt=0.05,
                                     dtmax=0.05, max_steps=10 ** 8, observer_x_up_thresh=0.3,
                                     observer_x_down_thresh=0.2)
        sim.set_repeat_ensemble(n)
        sim.seed_rng(seed)
        sim.enable_step_counter(True)
        sim.transient(t_span=(0.0, 100.0))  # compile and first launch, untimed
        sim.set_repeat_ensemble(n)
        sim.seed_rng(seed)
        sim.reset_step_counts()
        t0 = time.perf_counter()
        sim.transient(t_span=(0.0, 5000.0))
        k_tr = sim.last_kernel_ms()
        sim.set_tspan((0.0, 20000.0))
        sim.initialize_observer()
        k_init = sim.last_kernel_ms()
        sim.features(initialize_observer=False, fetch_results=False)
        k_feat = sim.last_kernel_ms()
        wall = time.perf_counter() - t0
        steps = int(sim.get_step_counts().sum())
        period = np.ravel(sim.get_observer_results().get_var_mean("period"))
        ok = period > 0
        res[prec] = {"member_steps": steps, "wall_s": wall, "member_steps_per_s": steps / wall,
                     "kernel_ms": {"transient": k_tr, "warmup": k_init, "features": k_feat},
                     "mean_period_ms": float(period[ok].mean()) if ok.any() else 0.0,
                     "period_cv": float(period[ok].std() / period[ok].mean()) if ok.any() else 0.0}
    res["fp32c_speedup_wall"] = res["double"]["wall_s"] / res["fp32c"]["wall_s"]
    return res


def lorenz_c3(n=1 << 20, t_end=500.0, n_traj=65536, store=1024):
    """Lorenz, 2^20 members (sigma 10, beta 8/3, rho in [0, 250], x0 ~ N(1, 0.1)),
    localmax on x over (0, 500), dopri5 and bs23 at rtol 1e-6 atol 1e-8, FP64
    and fp32c: kernel time and algorithmic FLOP rate of the features launch,
    wall time of a repeated features() call with the feature readback (the
    first call, which loads the program and allocates, is reported apart),
    and every fp32c feature's error against FP64 on the rho < 13 members;
    then the trajectory of the first 65,536 members, 1024 samples
    each (nout 1), float state, with its HBM write rate, and the same number of
    bytes from all 2^20 members at 64 samples each."""
    import clode_amd as clode

    rng = np.random.default_rng(20240601)
    x0 = rng.normal(1.0, 0.1, size=(n, 3))
    rho = np.linspace(0.0, 250.0, n)
    res = {}
    common = dict(src_file=os.path.join(MODELS, "lorenz.cl"), variables={"x": 1.0, "y": 1.0, "z": 1.0},
                  parameters={"sigma": 10.0, "rho": 28.0, "beta": 8.0 / 3.0}, dt=0.01, dtmax=1.0, abstol=1e-8,
                  reltol=1e-6, max_steps=100000000, t_span=(0.0, t_end))
    from clode_amd.perfmodel import PEAK_TFLOPS, launch_flops

    for stepper in (clode.Stepper.dormand_prince, clode.Stepper.bs23):
        r = {}
        workload = f"lorenz_2p20_{stepper.value}_localmax"
        for prec in ("double", "fp32c"):
            sim = clode.FeatureSimulator(observer=clode.Observer.local_max, stepper=stepper, single_precision=prec,
                                         **common)
            sim.set_ensemble(variables=x0, parameters={"rho": rho})
            # first call: device buffers, program and variant loads; untimed
            # for the rate, its features are the ones compared below
            t0 = time.perf_counter()
            first = sim.features(update_x0=False, fetch_results=True).to_ndarray()
            first_wall = time.perf_counter() - t0
            sim.enable_step_counter(True)
            t0 = time.perf_counter()
            out = sim.features(initialize_observer=True, update_x0=False, fetch_results=True)
            wall = time.perf_counter() - t0
            steps, trials = int(sim.get_step_counts().sum()), int(sim.get_trial_counts().sum())
            ms = sim.last_kernel_ms()
            fl = launch_flops(workload, "features", steps, trials)
            tf = fl / (ms / 1e3) / 1e12
            peak = PEAK_TFLOPS["f64"] if prec == "double" else PEAK_TFLOPS["f32"]
            r[prec] = {"kernel_ms": ms, "wall_s": wall, "first_call_wall_s": first_wall, "member_steps": steps,
                       "trials": trials, "member_steps_per_s": steps / wall, "algorithmic_tflops": tf,
                       "frac_of_peak": tf / peak, "peak_tflops": peak,
                       "occupancy_variant_waves_per_simd": sim.last_launch_variant()}
            r[prec + "_F"] = first
            r["names"] = out.get_feature_names()
        # fp32c feature error on the members without chaotic transients
        # (rho < 13), against FP64 at the same tolerance; the yardstick is the
        # method's own truncation error there, FP64 at rtol 1e-8 against 1e-6
        calm = rho < 13.0
        fine_sim = clode.FeatureSimulator(observer=clode.Observer.local_max, stepper=stepper, single_precision="double",
                                          **dict(common, abstol=1e-10, reltol=1e-8))
        fine_sim.set_ensemble(variables=x0[calm], parameters={"rho": rho[calm]})
        fine = fine_sim.features(fetch_results=True).to_ndarray()
        r["fp32c_feature_error_rho_lt_13"] = feature_error(r["fp32c_F"][calm], r["double_F"][calm], r["names"], fine)
        r["feature_error_members"] = int(calm.sum())
        for k in ("fp32c_F", "double_F", "names"):
            r.pop(k)
        r["fp32c_speedup_kernel"] = r["double"]["kernel_ms"] / r["fp32c"]["kernel_ms"]
        r["fp32c_speedup_wall"] = r["double"]["wall_s"] / r["fp32c"]["wall_s"]
        res[stepper.value] = r
    sim = clode.TrajectorySimulator(stepper=clode.Stepper.dormand_prince, single_precision="fp32c",
                                    max_store=store, nout=1, **common)
    sim.set_ensemble(variables=x0[:n_traj], parameters={"rho": rho[:n_traj]})
    sim.trajectory(fetch_results=False, update_x0=False)  # first launch (program load), untimed
    sim.trajectory(fetch_results=False, update_x0=False)
    ms = sim.last_kernel_ms()
    ns = np.array(sim._integrator.get_n_stored(), dtype=np.int64)
    samples = int(np.minimum(ns, store).sum())
    per_sample = 8 + 2 * 3 * 4  # double time + x, dx (float)
    gbs = samples * per_sample / (ms / 1e3) / 1e9
    res["trajectory_65536x1024_fp32c"] = {"kernel_ms": ms, "samples": samples, "bytes_per_sample": per_sample,
                                          "write_GBps": gbs, "frac_of_hbm_peak": gbs / 8000.0}
    # the same bytes from all 2^20 members x 64 samples: 16 waves of members
    # per SIMD instead of one, so the write stream is no longer paced by a
    # single wave's step latency
    sim = clode.TrajectorySimulator(stepper=clode.Stepper.dormand_prince, single_precision="fp32c",
                                    max_store=store // 16, nout=1, **common)
    sim.set_ensemble(variables=x0, parameters={"rho": rho})
    sim.trajectory(fetch_results=False, update_x0=False)
    sim.trajectory(fetch_results=False, update_x0=False)
    ms = sim.last_kernel_ms()
    ns = np.array(sim._integrator.get_n_stored(), dtype=np.int64)
    samples = int(np.minimum(ns, store // 16).sum())
    gbs = samples * per_sample / (ms / 1e3) / 1e9
    res["trajectory_2p20x64_fp32c"] = {"kernel_ms": ms, "samples": samples, "bytes_per_sample": per_sample,
                                       "write_GBps": gbs, "frac_of_hbm_peak": gbs / 8000.0}
    # the practical ceiling of a write-only stream: torch's fill of the same
    # number of bytes (median of 5 timed fills), in a process of its own (the
    # torch wheel carries its own HIP runtime, which does not share a process
    # with the one this library is linked against)
    fill_script = (
        "import sys, torch\n"
        "buf = torch.empty(int(sys.argv[1]) // 4, dtype=torch.float32, device='cuda')\n"
        "ts = []\n"
        "for k in range(6):\n"
        "    e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)\n"
        "    e0.record(); buf.fill_(float(k)); e1.record(); torch.cuda.synchronize()\n"
        "    ts.append(e0.elapsed_time(e1))\n"
        "ts = sorted(ts[1:])\n"
        "print(buf.numel() * 4 / (ts[len(ts) // 2] / 1e3) / 1e9)\n")
    try:
        out = subprocess.run([sys.executable, "-c", fill_script, str(samples * per_sample)], capture_output=True,
                             text=True, timeout=300, check=True)
        fill = float(out.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, OSError, ValueError, IndexError) as e:
        res["fill_same_bytes_error"] = f"{type(e).__name__}: {e}"[:300]
        return res
    res["fill_same_bytes_GBps"] = fill
    for k in ("trajectory_65536x1024_fp32c", "trajectory_2p20x64_fp32c"):
        res[k]["frac_of_fill_rate"] = res[k]["write_GBps"] / fill
    return res


def mixed_c4(n_each=1 << 18, rounds=20, seed=7, grid_rounds=3):
    """Four systems in equal shares (2^20 members): Van der Pol (dopri5,
    thresh2), Lorenz (dopri5, localmax), Chay-Keizer (dopri5, thresh2),
    Ornstein-Uhlenbeck (theta 1, sigma 0.5, seuler dt 0.05, basicall), all
    fp32c, per-member spans (0, 500 exp(z)), z ~ N(0, 1); rounds x (transient,
    x0 <- xf, features, x0 <- xf) as one captured graph of batched work-queue
    launches with autotuned launch shape, against the same sequence run system
    by system with one-member-per-lane grid launches."""
    import clode_amd as clode

    rng = np.random.default_rng(seed)

    def build():
        sims = []
        common = dict(single_precision="fp32c", max_steps=10 ** 9, abstol=1e-8, reltol=1e-6, dt=0.01, dtmax=1.0)
        vdp = clode.FeatureSimulator(src_file=os.path.join(MODELS, "van_der_pol.cl"),
                                     variables={"x": 2.0, "y": 0.0}, parameters={"mu": 1.0},
                                     observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                     observer_x_up_thresh=0.3, observer_x_down_thresh=0.2, **common)
        vdp.set_ensemble(parameters={"mu": rng.uniform(0.1, 5.0, n_each)})
        lor = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lorenz.cl"),
                                     variables={"x": 1.0, "y": 1.0, "z": 1.0},
                                     parameters={"sigma": 10.0, "rho": 28.0, "beta": 8.0 / 3.0},
                                     observer=clode.Observer.local_max, stepper=clode.Stepper.dormand_prince, **common)
        lor.set_ensemble(variables=rng.normal(1.0, 0.1, (n_each, 3)), parameters={"rho": rng.uniform(0, 50, n_each)})
        ck = clode.FeatureSimulator(src_file=os.path.join(MODELS, "chay_keizer.cl"),
                                    variables={"v": -50.0, "n": 0.01, "c": 0.12},
                                    parameters={"gca": 1200.0, "gkca": 750.0, "kpmca": 0.1},
                                    observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                    observer_x_up_thresh=0.3, observer_x_down_thresh=0.2, **common)
        ck.set_ensemble(parameters={"gca": rng.uniform(550, 1050, n_each), "kpmca": rng.uniform(0.095, 0.155, n_each)})
        ou = clode.FeatureSimulator(src_file=os.path.join(MODELS, "ornstein_uhlenbeck.cl"), variables={"x": 0.0},
                                    parameters={"theta": 1.0, "mu": 0.0, "sigma": 0.5}, num_noise=1,
                                    observer=clode.Observer.basic_all_variables, stepper=clode.Stepper.stochastic_euler,
                                    **dict(common, dt=0.05, dtmax=0.05))
        ou.set_ensemble(parameters={"theta": np.ones(n_each)})
        sims = [vdp, lor, ck, ou]
        for k, s in enumerate(sims):
            s.set_member_tspan(0.0, 500.0 * np.exp(rng.standard_normal(n_each)))
            s.seed_rng(100 + k)
            s.enable_step_counter(True)
        return sims

    state = rng.bit_generator.state
    sims = build()
    batch = clode.BatchSimulator(sims)
    t0 = time.perf_counter()
    block, slc = batch.autotune(stride=16, max_steps=4096)
    tune_s = time.perf_counter() - t0
    rows, cached = batch.autotune_table()
    tune = {"block": block, "slice": slc, "tune_s": tune_s, "from_cache": cached}
    if len(rows):
        order = np.argsort(rows[:, 2])
        b0, b1 = rows[order[0]], rows[order[1]]
        tune["trial_ms"] = {f"{int(r[0])}x{int(r[1])}": [round(float(v), 3) for v in r[2:]] for r in rows[order]}
        # the winner's lead over the runner-up, against the spread of their trials
        tune["lead_ms"] = float(b1[2] - b0[2])
        tune["spread_ms"] = float(max(b0[4] - b0[3], b1[4] - b1[3]))
    for s in sims:
        s.reset_step_counts()
    t0 = time.perf_counter()
    batch.sweep(rounds, graph=True)
    graph_s = time.perf_counter() - t0
    steps = sum(int(s.get_step_counts().sum()) for s in sims)
    F_graph = [s.get_observer_results().to_ndarray() for s in sims]
    roof = c4_round_roofline(batch, sims)
    # the same sequence, system by system, one-member-per-lane grid launches,
    # over the first grid_rounds rounds (compared per round)
    rng.bit_generator.state = state
    sims2 = build()
    grid_rounds = min(rounds, grid_rounds)
    t0 = time.perf_counter()
    for _ in range(grid_rounds):
        for s in sims2:
            s.transient()
            s.features(fetch_results=False)
    grid_s = time.perf_counter() - t0
    same = None
    if grid_rounds == rounds:
        same = all(np.array_equal(a, b.get_observer_results().to_ndarray()) for a, b in zip(F_graph, sims2))
    return {"members": 4 * n_each, "rounds": rounds, "autotuned": tune,
            "graph_sweep_s": graph_s, "s_per_round": graph_s / rounds, "member_steps": steps,
            "member_steps_per_s": steps / graph_s, "grid_rounds": grid_rounds, "grid_s_per_round": grid_s / grid_rounds,
            "speedup_vs_grid_per_round": (grid_s / grid_rounds) / (graph_s / rounds),
            "identical_features": same, "round_roofline": roof}


def c4_round_roofline(batch, sims):
    """One more round of the mixed sweep, transient and features as separate
    batched launches, priced with the generated operation table: each
    system's device-counted steps and trials per phase times its per-step and
    per-rejected-trial FLOPs, over the two launches' device time.  (The
    observers continue, so the features launch has no warm-up pass.)"""
    from clode_amd.perfmodel import PEAK_TFLOPS, launch_flops

    names = ("mixed_vdp_dopri5_thresh2", "mixed_lorenz_dopri5_localmax", "mixed_chay_keizer_dopri5_thresh2",
             "mixed_ou_seuler_basicall")
    flops, ms, per = 0.0, 0.0, {}
    for phase, run in (("transient", lambda: batch.transient(update_x0=False)),
                       ("features", lambda: batch.features(update_x0=False))):
        for s in sims:
            s.reset_step_counts()
        run()
        ms += batch.last_kernel_ms()
        batch.shift_x0()
        for w, s in zip(names, sims):
            st, tr = int(s.get_step_counts().sum()), int(s.get_trial_counts().sum())
            f = launch_flops(w, phase, st, tr)
            flops += f
            per.setdefault(w, {})[phase] = {"member_steps": st, "trials": tr, "flops": f}
    tf = flops / (ms / 1e3) / 1e12
    return {"kernel_ms": ms, "algorithmic_flops": flops, "algorithmic_tflops": tf,
            "frac_of_fp32_peak": tf / PEAK_TFLOPS["f32"], "systems": per}


def surrogate_c5(precision, n=2048, held_out=False):
    """2048 x 2048 lactotroph feature map (gK in [0, 10] x gBK in [0, 2],
    dopri5 rtol 1e-6, transient 5e3 then thresh2 over 2e4): the plain map
    against the surrogate-screened map (2^18 integrated training samples,
    BF16-MFMA MLP trained with screened_map's defaults, screening, re-solve of
    the flagged cells).  screened_map's defaults were chosen on this map with
    training seed 1; with held_out, the same defaults are checked on data
    they were not tuned on: training seeds 2 and 3 on this map, and a second
    map over the shifted box gK in [2, 12] x gBK in [0.5, 2.5] (its own plain
    map as the reference)."""
    import clode_amd as clode
    from clode_amd.surrogate import compare_maps, integrator_solver, screened_map

    sim = clode.FeatureSimulator(src_file=os.path.join(MODELS, "lactotroph.cl"),
                                 variables={"v": -60.0, "n": 0.0, "f": 0.0, "c": 0.1},
                                 parameters={"gcal": 1.5, "gsk": 3.0, "gbk": 1.0, "gk": 2.0}, aux=["ica"],
                                 observer=clode.Observer.threshold_2, stepper=clode.Stepper.dormand_prince,
                                 single_precision=precision, dt=0.1, dtmax=1.0, abstol=1e-8, reltol=1e-6,
                                 max_steps=10 ** 8, observer_x_up_thresh=0.3, observer_x_down_thresh=0.2)
    solve = integrator_solver(sim, ("gk", "gbk"), ((0.0, 10.0), (0.0, 2.0)), 5000.0, 20000.0, "v")
    # compile and first launch, untimed; 2^18 points so that the variant used
    # for large ensembles (compiled on first use) is built here too
    solve(np.random.default_rng(0).random((1 << 18, 2)))
    gx, gy = np.meshgrid(np.arange(n) / (n - 1), np.arange(n) / (n - 1), indexing="ij")
    t0 = time.perf_counter()
    plain = solve(np.stack([gx.ravel(), gy.ravel()], axis=1))
    plain_s = time.perf_counter() - t0
    plain = {k: v.reshape(n, n) for k, v in plain.items()}
    m, _ = screened_map(solve, n, n, seed=1)
    bad = compare_maps({"period": m.period, "amp": m.amp, "cls": m.cls}, plain, tol=0.05)
    defaults = inspect.signature(screened_map).parameters
    checks = None
    if held_out:
        checks = {}
        for seed in (2, 3):
            ms, _ = screened_map(solve, n, n, seed=seed)
            checks[f"training_seed_{seed}"] = {
                "cells_disagreeing_frac": float(compare_maps({"period": ms.period, "amp": ms.amp, "cls": ms.cls},
                                                             plain, tol=0.05).mean()),
                "screened_total_s": ms.timings["total_s"], "cells_integrated_frac": float(ms.solved.mean())}
        box = ((2.0, 12.0), (0.5, 2.5))
        solve2 = integrator_solver(sim, ("gk", "gbk"), box, 5000.0, 20000.0, "v")
        t0 = time.perf_counter()
        plain2 = solve2(np.stack([gx.ravel(), gy.ravel()], axis=1))
        plain2_s = time.perf_counter() - t0
        plain2 = {k: v.reshape(n, n) for k, v in plain2.items()}
        m2, _ = screened_map(solve2, n, n, seed=1)
        bad2 = compare_maps({"period": m2.period, "amp": m2.amp, "cls": m2.cls}, plain2, tol=0.05)
        checks["shifted_box_gk_2_12_gbk_0.5_2.5"] = {
            "plain_map_s": plain2_s, "screened_total_s": m2.timings["total_s"],
            "speedup_total": plain2_s / m2.timings["total_s"], "cells_disagreeing_frac": float(bad2.mean()),
            "cells_integrated_frac": float(m2.solved.mean())}
    return {"grid": f"{n}x{n}", "training_samples": defaults["n_train"].default,
            "epochs": defaults["epochs"].default, "dilation": defaults["dilate"].default,
            "plain_map_s": plain_s, "screened_total_s": m.timings["total_s"],
            "speedup_total": plain_s / m.timings["total_s"],
            "speedup_screen_and_resolve": plain_s / (m.timings["screen_s"] + m.timings["resolve_s"]),
            "timings": m.timings, "cells_integrated_frac": float(m.solved.mean()),
            "cells_disagreeing_frac": float(bad.mean()), "disagreement_rule": "class differs, or period or "
            "amplitude differs by more than 5 % relative to the plain map", "final_train_loss":
            float(m.train_loss[-1]) if m.train_loss is not None else None, "held_out": checks}


def profiled_value(key, profile="features_lactotroph_256x256_dopri5_thresh2_fp64.md"):
    """A figure measured by rocprofv3 and committed in a profile summary
    (a line 'key value' under profiles/), or None."""
    import re

    try:
        text = open(os.path.join(REPO, "profiles", profile)).read()
    except OSError:
        return None
    m = re.search(r"^%s\s+([0-9.eE+-]+)\s*$" % re.escape(key), text, re.M)
    return float(m.group(1)) if m else None


def profiled_traffic(kernel="clode_features"):
    """Scaled HBM bytes per launch of the kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload."""
    import re

    path = os.path.join(REPO, "profiles", "features_lactotroph_256x256_dopri5_thresh2_fp64.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\s+%s\s+(\d+)" % kernel, text, re.M)
    return int(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--cpu-members", type=int, default=1024)
    ap.add_argument("--extras", default="fp32c,seq,replay,c1,noise,c3,c4,c5",
                    help="comma list of extra workloads: fp32c, seq (one-launch sequence), replay (features "
                         "replay), c1, noise, c3, c4, c5 "
                         "(or 'none'); --no-extra "
                         "drops them all")
    ap.add_argument("--c4-rounds", type=int, default=8)
    ap.add_argument("--mlp-kernels", action="store_true",
                    help="run only the surrogate kernels (one training epoch over 65,536 synthetic samples and one "
                         "2048 x 2048 screening pass), for profiling")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 256 MiB device copy kernel (FETCH_SIZE/WRITE_SIZE calibration)")
    ap.add_argument("--calibrate-fma", action="store_true",
                    help="run only the FP32/FP64 FMA throughput probe (achievable vector-ALU peaks)")
    args = ap.parse_args()
    if args.calibrate_fma:
        import subprocess

        probe = os.path.join(REPO, "clode_amd", "_build", "valu_rate")
        if not os.path.exists(probe):
            sys.exit("bench.py: the FMA probe is not built (run __graft_entry__.build())")
        sys.exit(subprocess.run([probe]).returncode)
    if args.calibrate_copy:
        import torch

        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3}))
        return
    if args.mlp_kernels:
        from clode_amd.surrogate import MLPSurrogate

        rng = np.random.default_rng(0)
        U = rng.random((65536, 2))
        T = np.stack([np.sin(3 * U[:, 0]), U[:, 0] * U[:, 1]], axis=1)
        cls = np.clip((8 * U[:, 0]).astype(int), 0, 15)
        net = MLPSurrogate(seed=1)
        net.train(U, T, cls, epochs=1, batch=2048)  # first use of every kernel, untimed
        hist = net.train(U, T, cls, epochs=1, batch=2048)
        train_ms = net.last_ms()
        net.screen_grid(2048, 2048)
        print(json.dumps({"train_epoch_ms": train_ms, "train_loss": float(hist[-1]), "screen_ms": net.last_ms(),
                          "screen_tflops": 4194304 * 2 * 238720 / (net.last_ms() / 1e3) / 1e12}))
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU; multi-GPU runs are not supported "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    from clode_amd.perfmodel import PEAK_TFLOPS, phase_counts

    r64 = run_lactotroph("double", args.steps, args.warmup)
    print(f"# fp64: {r64['member_steps_per_s']:.4e} member-steps/s, {r64['ms_per_step']:.1f} ms/step, "
          f"kernels {r64['kernel_ms']}", file=sys.stderr, flush=True)
    achieved, feat_flops = kernel_tflops(r64, "features")
    pc = phase_counts("lactotroph_256x256_dopri5_thresh2", "features")
    executed = profiled_value("executed_fp64_flops_per_launch")
    one_wave = profiled_value("fp64_fma_tflops_one_wave_per_simd", "valu_fma_rate.md")
    roofline = {"bound": "valu-fp64", "achieved": achieved, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS["f64"], "traffic": profiled_traffic(),
                "kernel": "clode_features (thresh2, dopri5, FP64)",
                "algorithmic_flops_per_launch": feat_flops,
                "flops_per_member_step": pc["step"], "flops_per_rejected_trial": pc["rejected_trial"],
                "transcendentals_per_member_step": pc["transcendentals"],
                "member_steps_per_launch": r64["kernel_member_steps"]["features"],
                "trials_per_launch": r64["kernel_trials"]["features"], "launch_ms": r64["kernel_ms"]["features"],
                "flop_table": "clode_amd/perfmodel.py (generated by oracle/gen_perfmodel.py)"}
    if executed:
        ex = executed / (r64["kernel_ms"]["features"] / 1e3) / 1e12
        roofline["executed_tflops"] = ex
        roofline["executed_frac"] = ex / PEAK_TFLOPS["f64"]
    if one_wave:
        # the launch has exactly one wave per SIMD (65,536 members, 1,024 SIMDs)
        roofline["one_wave_per_simd_fma_ceiling_tflops"] = one_wave
        roofline["frac_of_one_wave_ceiling"] = achieved / one_wave
        if executed:
            roofline["executed_frac_of_one_wave_ceiling"] = roofline["executed_tflops"] / one_wave

    extra = {}
    wanted = set() if args.no_extra or args.extras == "none" else set(args.extras.split(","))

    def run_extra(name, fn):
        t0 = time.perf_counter()
        try:
            extra[name] = fn()
        except Exception as e:  # an extra figure must not cost the headline line
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
        print(f"# {name}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    r32 = None
    if "fp32c" in wanted:
        def fp32c_extra():
            nonlocal r32
            r32 = run_lactotroph("fp32c", args.steps, args.warmup)
            print(f"# fp32c: {r32['member_steps_per_s']:.4e} member-steps/s, {r32['ms_per_step']:.1f} ms/step",
                  file=sys.stderr, flush=True)
            a32, _ = kernel_tflops(r32, "features")
            return {
                "member_steps_per_s": r32["member_steps_per_s"], "ms_per_step": r32["ms_per_step"],
                "kernel_ms": r32["kernel_ms"], "speedup_vs_fp64_wall": r64["ms_per_step"] / r32["ms_per_step"],
                "features_kernel_tflops": a32, "features_kernel_frac_of_fp32_peak": a32 / PEAK_TFLOPS["f32"],
                "feature_error_vs_fp64": feature_error(
                    lactotroph_step_features("fp32c", r64["x0_step"]),
                    lactotroph_step_features("double", r64["x0_step"]), r64["feature_names"],
                    lactotroph_step_features("double", r64["x0_step"], 1e-8, 1e-10)),
                "feature_error_members": "all 65,536, one step from the same initial states (fresh simulators)",
            }

        run_extra("fp32c", fp32c_extra)
    if "seq" in wanted:
        run_extra("one_launch_sequence_fp64", lambda: one_launch_sequence("double", args.steps, r64))
        if r32 is not None:
            run_extra("one_launch_sequence_fp32c", lambda: one_launch_sequence("fp32c", args.steps, r32))
    if "replay" in wanted:
        run_extra("features_replay_fp64", lambda: replay_step("double", args.steps, r64))
        if r32 is not None:
            run_extra("features_replay_fp32c", lambda: replay_step("fp32c", args.steps, r32))
    if "c1" in wanted:
        run_extra("vdp_4096_rk4_fp64", vdp_c1)
    if "noise" in wanted:
        run_extra("lactotroph_noise_65536_seuler", noise_c2)
    if "c3" in wanted:
        run_extra("lorenz_2p20_localmax", lorenz_c3)
    if "c4" in wanted:
        run_extra("mixed_2p20_persistent_sweep", lambda: mixed_c4(rounds=args.c4_rounds))
    if "c5" in wanted:
        run_extra("surrogate_2048_fp64", lambda: surrogate_c5("double", held_out=True))
        run_extra("surrogate_2048_fp32c", lambda: surrogate_c5("fp32c"))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_members)
        print(f"# cpu baseline: {cpu['value']:.4e} member-steps/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    # the second half of the metric: the 2048^2 map's wall time, plain and
    # screened, with the share of cells that disagree with the plain map
    maps = {}
    for prec in ("fp64", "fp32c"):
        m = extra.get(f"surrogate_2048_{prec}")
        if m and "error" not in m:
            maps[prec] = {"plain_map_s": m["plain_map_s"], "screened_map_s": m["screened_total_s"],
                          "speedup": m["speedup_total"], "cells_disagreeing_frac": m["cells_disagreeing_frac"]}
    line = {
        "metric": "member-steps/s + 2048\u00b2 feature-map wall time at \u22641% feature error vs plain",
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
        "map_2048": maps or None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
