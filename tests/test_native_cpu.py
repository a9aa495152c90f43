"""CPU-side checks of the native library and its boundary (no GPU needed):
exported C ABI, pybind11 module surface and defaults (clode/cpp/CLODEpython.cpp),
observer layouts and feature names, hipRTC compilation of every program
family for gfx950, and the runtime's argument validation."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "clode_amd", "libclode_hip.so")
HEADER = os.path.join(REPO, "include", "clode_hip.h")
MODELS = os.path.join(REPO, "clode_amd", "models")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(clode_\w+)\(", text, re.M)))


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_surrogate_abi_exports_every_declared_symbol():
    text = open(os.path.join(REPO, "include", "clode_mlp.h")).read()
    names = sorted(set(re.findall(r"^(?:int|const char \*)\s*(clode_mlp_\w+)\(", text, re.M)))
    assert len(names) >= 12
    lib = ctypes.CDLL(os.path.join(REPO, "clode_amd", "libclode_mlp.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # no device here: creation fails cleanly with a message
    h = ctypes.c_void_p()
    lib.clode_mlp_last_error.restype = ctypes.c_char_p
    assert lib.clode_mlp_create(ctypes.byref(h), 0, ctypes.c_ulonglong(0)) == 2
    assert b"device" in lib.clode_mlp_last_error()
    assert lib.clode_mlp_destroy(None) == 0


def test_c_abi_without_device():
    lib = ctypes.CDLL(LIB)
    n = ctypes.c_int(-1)
    assert lib.clode_device_count(ctypes.byref(n)) == 0 and n.value == 0
    lib.clode_last_error.restype = ctypes.c_char_p
    assert lib.clode_destroy(None) == 0
    assert lib.clode_transient(None) == 1  # CLODE_ERR_INVALID on a null handle
    assert b"null" in lib.clode_last_error()


class _ProblemInfo(ctypes.Structure):
    _fields_ = [("src_file", ctypes.c_char_p), ("num_var", ctypes.c_int), ("num_par", ctypes.c_int),
                ("num_aux", ctypes.c_int), ("num_noise", ctypes.c_int), ("var_names", ctypes.c_char_p),
                ("par_names", ctypes.c_char_p), ("aux_names", ctypes.c_char_p)]


def test_c_abi_null_pointers_are_invalid_arguments():
    """Null output pointers and null input arrays with a nonzero size are
    refused with CLODE_ERR_INVALID and a message, never dereferenced."""
    lib = ctypes.CDLL(LIB)
    lib.clode_last_error.restype = ctypes.c_char_p
    assert lib.clode_device_count(None) == 1
    assert lib.clode_get_log_level(None) == 1
    pi = _ProblemInfo(os.path.join(MODELS, "van_der_pol.cl").encode(), 2, 1, 0, 0, b"x\ny", b"mu", None)
    sim = ctypes.c_void_p()
    assert lib.clode_create(ctypes.byref(sim), 0, ctypes.byref(pi), b"rk4", None, None, 2, 0, None) == 0
    try:
        for fn, args in [("clode_get_n_pts", [None]), ("clode_get_tspan", [None]),
                         ("clode_get_solver_params", [None]), ("clode_get_precision", [None]),
                         ("clode_get_problem_info", [None, None, None, None]),
                         ("clode_is_observer_initialized", [None]), ("clode_last_kernel_ms", [None]),
                         ("clode_set_x0", [None, ctypes.c_size_t(4)]),
                         ("clode_set_pars", [None, ctypes.c_size_t(2)]),
                         ("clode_set_tspan", [None, ctypes.c_size_t(2)]),
                         ("clode_set_problem_data", [None, ctypes.c_size_t(2), None, ctypes.c_size_t(1)])]:
            assert getattr(lib, fn)(sim, *args) == 1, fn
            assert b"null" in lib.clode_last_error(), fn
        n = ctypes.c_int(-1)
        assert lib.clode_get_n_pts(sim, ctypes.byref(n)) == 0 and n.value == 0
    finally:
        assert lib.clode_destroy(sim) == 0


def test_pybind_defaults_match_reference():
    from clode_amd.cpp import clode_cpp_wrapper as w

    sp = w.SolverParams()
    assert (sp.dt, sp.dtmax, sp.abstol, sp.reltol, sp.max_steps, sp.max_store, sp.nout) == \
           (0.1, 0.5, 1e-6, 1e-3, 1000000, 1000000, 1)
    op = w.ObserverParams()
    assert (op.e_var_ix, op.f_var_ix, op.max_event_count, op.max_event_timestamps) == (0, 0, 100, 0)
    assert (op.min_amp, op.min_imi, op.nhood_radius, op.x_up_threshold, op.x_down_threshold) == \
           (0.0, 0.0, 0.05, 0.2, 0.2)
    pi = w.ProblemInfo("f.cl", ["x", "y"], ["a"])
    assert pi.num_noise == 1 and pi.num_var == 2 and pi.num_par == 1 and pi.num_aux == 0
    pi.vars = ["u"]
    assert pi.num_var == 1
    for cls in ("SimulatorBase", "FeatureSimulatorBase", "TrajectorySimulatorBase", "OpenCLResource",
                "CLDeviceType", "CLVendor", "DeviceInfo", "PlatformInfo", "LogLevel", "LoggerSingleton"):
        assert hasattr(w, cls)
    for fn in ("query_opencl", "_print_opencl", "get_logger"):
        assert hasattr(w, fn)
    # get_logger returns the LoggerSingleton instance (CLODEpython.cpp:87-94)
    logger = w.get_logger()
    assert type(logger).__name__ == "LoggerSingleton"
    for fn in ("set_log_level", "set_log_pattern", "get_log_level"):
        assert hasattr(logger, fn)


@pytest.mark.parametrize("observer,nvar,naux,nstore", [
    ("thresh2", 4, 1, 0), ("thresh2", 2, 0, 3), ("localmax", 3, 0, 2), ("localmax", 3, 1, 0),
    ("nhood2", 2, 0, 1), ("nhood2", 4, 2, 0), ("basic", 3, 1, 0), ("basicall", 4, 2, 0), ("basicall", 2, 0, 1),
    ("nhood1", 3, 0, 0), ("nhood1", 4, 1, 2)])
def test_feature_names_and_counts(observer, nvar, naux, nstore):
    from clode_amd.cpp import clode_cpp_wrapper as w

    model = {2: "van_der_pol", 3: "lorenz", 4: "lactotroph"}[nvar]
    src = os.path.join(MODELS, model + ".cl")
    vars_ = [f"v{j}" for j in range(nvar)]
    pars = ["p0", "p1", "p2", "p3"][: {2: 1, 3: 3, 4: 4}[nvar]]
    aux = [f"a{j}" for j in range(naux)]
    op = w.ObserverParams(max_event_timestamps=nstore)
    s = w.FeatureSimulatorBase(w.ProblemInfo(src, vars_, pars, aux, 0), "rk4", observer, op, False,
                               w.OpenCLResource(), "")
    names = s.get_feature_names()
    expected = {"thresh2": 18 + 5 * nvar + 3 * naux + 2 * nstore + 5,
                "localmax": 6 + 5 * nvar + 3 * naux + 4 * nstore + 2,
                "nhood2": 6 + 7 * nvar + 3 * naux + nstore + 5, "basic": 6,
                "basicall": 5 * nvar + 3 * naux + 1, "nhood1": 6 + 5 * nvar + 3 * naux + 5}[observer]
    assert s.get_n_features() == expected == len(names) == len(set(names))
    assert names[-1 if observer in ("localmax", "basic", "basicall") else -4] == "step count"
    if observer == "basic":  # the feature variable only (f_var_ix 0)
        assert names[:5] == ["max v0", "min v0", "mean v0", "max dv0/dt", "min dv0/dt"]
        return
    assert f"max d{vars_[-1]}/dt" in names
    if observer == "nhood1":
        assert names[-5] == "period count"
    if nstore and observer in ("thresh2", "localmax", "nhood2"):
        key = {"thresh2": "up event time 0", "localmax": "localmax event evar 0", "nhood2": "nhood event time 0"}
        assert key[observer] in names


@pytest.mark.parametrize("obs", [1, 2, 3, 4, 5, 6])
def test_observer_state_layout_matches_host(tmp_path, obs):
    """Compile the device observer header on the host and count the
    persistent fields it spills, against clode_layout.h (host allocation)."""
    exe = tmp_path / "probe"
    for nv, na in ((1, 0), (2, 1), (4, 3)):
        subprocess.run(["g++", "-std=c++17", "-DPREC=2", f"-DCLODE_OBSERVER={obs}", f"-DN_VAR={nv}",
                        f"-DN_AUX={na}", "-DN_STORE_EVENTS=2", os.path.join(REPO, "tests", "layout_probe.cpp"),
                        "-o", str(exe)], check=True, cwd=os.path.join(REPO, "tests"))
        r = subprocess.run([str(exe)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("kind,stepper,observer,prec", [
    ("features", "rk4", "thresh2", "double"), ("features", "dopri5", "thresh2", "fp32c"),
    ("features", "bs23", "localmax", "single"), ("features", "seuler", "nhood2", "double"),
    ("features", "heun", "nhood2", "fp32c"), ("features", "euler", "localmax", "fp32c"),
    ("features", "rk4", "basic", "double"), ("features", "dopri5", "basicall", "fp32c"),
    ("features", "bs23", "nhood1", "single"), ("features", "rk4", "nhood1", "fp32c"),
    ("trajectory", "dopri5", None, "single"), ("trajectory", "rk4", None, "fp32c"),
    ("base", "bs23", None, "double")])
def test_hiprtc_compiles_for_gfx950(kind, stepper, observer, prec):
    from clode_amd.cpp import clode_cpp_wrapper as w

    pi = w.ProblemInfo(os.path.join(MODELS, "lactotroph_noise.cl"), ["v", "n", "f", "c"],
                       ["gcal", "gsk", "gbk", "sigma"], ["ica"], 1)
    dev = w.OpenCLResource()
    if kind == "features":
        s = w.FeatureSimulatorBase(pi, stepper, observer, w.ObserverParams(max_event_timestamps=2), prec, dev, "")
    elif kind == "trajectory":
        s = w.TrajectorySimulatorBase(pi, stepper, prec, dev, "")
    else:
        s = w.SimulatorBase(pi, stepper, prec, dev, "")
    s.compile()
    prog = s.get_program_string()
    assert "--offload-arch=gfx950" in prog and "-packed-fp32-ops" in prog
    assert f"#define CLODE_PREC {dict(single=1, double=2, fp32c=3)[prec]}" in prog
    # only fp32c relaxes float semantics; single and double programs must
    # round exactly as the oracle does
    relaxed = ("-fgpu-flush-denormals-to-zero", "-freciprocal-math", "-fapprox-func",
               "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fno-honor-nans", "-mno-amdgpu-ieee")
    for flag in relaxed:
        assert (flag in prog) == (prec == "fp32c"), flag
    assert ("-fhip-fp32-correctly-rounded-divide-sqrt" in prog) == (prec != "fp32c")


@pytest.mark.parametrize("stepper,observer,prec", [("dopri5", "thresh2", "double"), ("bs23", "nhood2", "single"),
                                                    ("rk4", "thresh2", "fp32c")])
def test_replay_program_compiles_for_gfx950(stepper, observer, prec):
    """With a replay budget, compile() also builds the program holding the
    warm-up record and features replay kernels (on the CPU, for gfx950); the
    replay state reads (0 records, not replayed) until a warm-up runs, and
    the C ABI refuses null outputs."""
    from clode_amd.cpp import clode_cpp_wrapper as w

    pi = w.ProblemInfo(os.path.join(MODELS, "lactotroph.cl"), ["v", "n", "f", "c"], ["gcal", "gsk", "gbk", "gk"],
                       ["ica"], 0)
    s = w.FeatureSimulatorBase(pi, stepper, observer, w.ObserverParams(), prec, w.OpenCLResource(), "")
    assert s.get_replay_info() == (0, False)
    s.set_replay_budget(1 << 30)
    s.compile()
    s.set_replay_budget(0)
    assert s.get_replay_info() == (0, False)
    lib = ctypes.CDLL(LIB)
    assert lib.clode_set_replay_budget(None, ctypes.c_ulonglong(1)) == 1
    assert lib.clode_get_replay_info(None, None, None) == 1


_COMPILE_INTO = """
import sys
from clode_amd.cpp import clode_cpp_wrapper as w
pi = w.ProblemInfo(sys.argv[1], ["x", "y"], ["mu"], [], 0)
w.SimulatorBase(pi, "rk4", "double", w.OpenCLResource(), sys.argv[2]).compile()
"""


def test_code_object_cache_checksum(tmp_path):
    """The on-disk code-object cache: a build writes one file (magic, key,
    size, checksum, code); a fresh process reuses it untouched; a file with a
    damaged body is ignored, rebuilt and rewritten."""
    src = os.path.join(MODELS, "van_der_pol.cl")

    def compile_in_fresh_process():
        subprocess.run(["python3", "-c", _COMPILE_INTO, src, str(tmp_path)], check=True, cwd=REPO)

    compile_in_fresh_process()
    files = sorted(tmp_path.glob("clode_*.co"))
    assert len(files) == 1
    path = files[0]
    data = path.read_bytes()
    assert data[:8] == b"CLODECO2"
    n = int.from_bytes(data[16:24], "little")
    assert len(data) == 32 + n
    h = 1469598103934665603
    for c in data[32:]:
        h = ((h ^ c) * 1099511628211) & (2**64 - 1)
    assert int.from_bytes(data[24:32], "little") == h
    stamp = path.stat().st_mtime_ns
    compile_in_fresh_process()
    assert path.stat().st_mtime_ns == stamp  # intact: reused, not rewritten
    damaged = bytearray(data)
    damaged[32 + n // 2] ^= 0xFF
    path.write_bytes(bytes(damaged))
    compile_in_fresh_process()
    rebuilt = path.read_bytes()
    assert rebuilt[:32] == data[:32] and rebuilt == data
    # a truncated file, and a header whose size field claims more than the
    # file holds (2^40 bytes: checked against the file length before any
    # allocation), are ignored and rebuilt as well
    path.write_bytes(data[: 32 + n // 3])
    compile_in_fresh_process()
    assert path.read_bytes() == data
    huge = bytearray(data)
    huge[16:24] = (1 << 40).to_bytes(8, "little")
    path.write_bytes(bytes(huge))
    compile_in_fresh_process()
    assert path.read_bytes() == data
    assert not list(tmp_path.glob("*.tmp*"))


def test_code_object_cache_key_covers_the_toolchain(tmp_path):
    """The cache key includes the toolchain identity (HIP release, runtime
    version, the loaded hipRTC library and CLODE_TOOLCHAIN_ID): the same
    program under another identity text gets another cache file, and the
    original file is left alone."""
    src = os.path.join(MODELS, "van_der_pol.cl")

    def compile_with(extra_env):
        env = dict(os.environ)
        env.pop("CLODE_TOOLCHAIN_ID", None)
        env.update(extra_env)
        subprocess.run(["python3", "-c", _COMPILE_INTO, src, str(tmp_path)], check=True, cwd=REPO, env=env)

    compile_with({})
    first = sorted(p.name for p in tmp_path.glob("clode_*.co"))
    compile_with({})
    assert sorted(p.name for p in tmp_path.glob("clode_*.co")) == first
    compile_with({"CLODE_TOOLCHAIN_ID": "another compiler"})
    both = sorted(p.name for p in tmp_path.glob("clode_*.co"))
    assert len(first) == 1 and len(both) == 2 and first[0] in both


def test_unknown_names_and_bad_input():
    from clode_amd.cpp import clode_cpp_wrapper as w

    pi = w.ProblemInfo(os.path.join(MODELS, "van_der_pol.cl"), ["x", "y"], ["mu"], [], 0)
    with pytest.raises(ValueError):
        w.SimulatorBase(pi, "rk5", False, w.OpenCLResource(), "")
    with pytest.raises(ValueError):
        w.FeatureSimulatorBase(pi, "rk4", "nosuch", w.ObserverParams(), False, w.OpenCLResource(), "")
    with pytest.raises(ValueError):
        w.FeatureSimulatorBase(pi, "rk4", "thresh2", w.ObserverParams(e_var_ix=5), False, w.OpenCLResource(), "")
    with pytest.raises(ValueError):
        w.SimulatorBase(w.ProblemInfo("/nonexistent.cl", ["x"], []), "rk4", False, w.OpenCLResource(), "")
    s = w.SimulatorBase(pi, "rk4", False, w.OpenCLResource(), "")
    s.set_stepper("nosuch")  # reference: warning, stepper unchanged
    assert "#define CLODE_STEPPER 3" in s.get_program_string()
    with pytest.raises(ValueError, match="at least one member"):
        s.set_problem_data([], [])  # empty ensemble, refused before any device work
    bad = w.SolverParams()
    bad.nout = 0
    with pytest.raises(ValueError, match="nout"):
        s.set_solver_params(bad)
    with pytest.raises(RuntimeError):
        s.set_problem_data([1.0, 2.0], [1.0])  # no device here


def test_runtime_validation():
    """Reference test/test_runtime.py: invalid device-selection combinations."""
    from clode_amd.runtime import CLDeviceType, CLVendor, initialize_runtime

    with pytest.raises(ValueError):
        initialize_runtime(CLDeviceType.DEVICE_TYPE_GPU, None, 0, 0, None)
    with pytest.raises(ValueError):
        initialize_runtime(None, CLVendor.VENDOR_AMD, 0, 0, None)
    with pytest.raises(ValueError):
        initialize_runtime(None, None, 0, 0, [0])
    with pytest.raises(ValueError):
        initialize_runtime(None, None, 0, None, None)
    with pytest.raises(ValueError):
        initialize_runtime(None, None, None, 0, None)
    with pytest.raises(ValueError):
        initialize_runtime(None, None, None, None, [0])
    assert initialize_runtime(None, None, 0, 1, None).device == 1
    assert initialize_runtime(None, None, None, None, None).device == 0


def test_batch_program_compiles_for_gfx950():
    """Four different systems, steppers and observers in one batched program
    (one namespace per system and dispatching work-queue kernels)."""
    from clode_amd.cpp import clode_cpp_wrapper as w

    dev = w.OpenCLResource()
    op = w.ObserverParams(max_event_timestamps=1)
    sims = [
        w.FeatureSimulatorBase(w.ProblemInfo(os.path.join(MODELS, "van_der_pol.cl"), ["x", "y"], ["mu"], [], 0),
                               "rk4", "thresh2", op, "fp32c", dev, ""),
        w.FeatureSimulatorBase(w.ProblemInfo(os.path.join(MODELS, "lorenz.cl"), ["x", "y", "z"], ["s", "r", "b"],
                                             [], 0), "dopri5", "localmax", op, "fp32c", dev, ""),
        w.FeatureSimulatorBase(w.ProblemInfo(os.path.join(MODELS, "chay_keizer.cl"), ["v", "n", "c"],
                                             ["gca", "gkca", "kpmca"], [], 0), "bs23", "nhood2", op, "fp32c", dev, ""),
        w.FeatureSimulatorBase(w.ProblemInfo(os.path.join(MODELS, "ornstein_uhlenbeck.cl"), ["x"],
                                             ["theta", "mu", "sigma"], [], 1), "seuler", "basicall", op, "fp32c",
                               dev, ""),
    ]
    w.compile_batch(sims)
    mixed = [sims[0], w.FeatureSimulatorBase(
        w.ProblemInfo(os.path.join(MODELS, "decay.cl"), ["x"], ["lambda"], [], 0), "rk4", "basic", op, "double",
        dev, "")]
    with pytest.raises(ValueError):
        w.compile_batch(mixed)  # precisions differ
