"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/oracle_driver.cpp runs every kind (base, features,
trajectory) x stepper x observer x precision, with max_store 0, 1 and 7,
observer event lists, and per-member time spans, on two models (a plain ODE
and a stochastic one).  Any out-of-bounds access or undefined operation in
the oracle fails this test.  Host code only (clang++ from ROCm)."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "bin", "clang++")
OMP_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "lib")

# model, nVar, nPar, nAux, nWiener, x0 scale, parameter scale, span end
CASES = [
    ("van_der_pol", 2, 1, 0, 0, 1.0, 1.0, 20.0),
    ("lactotroph_noise", 4, 4, 1, 1, -100.0, 5.0, 50.0),
]


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang++ not found")
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_under_sanitizers(tmp_path, case):
    model, nv, npar, na, nw, xs, ps, tend = case
    exe = tmp_path / f"oracle_{model}"
    cmd = [CLANG, "-g", "-O1", "-std=c++17", "-ffp-contract=on", "-mfma", "-fopenmp",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-Wno-unused-parameter",
           f'-DCLODE_ORACLE_RHS="{os.path.join(REPO, "clode_amd", "models", model + ".cl")}"',
           f"-DNV={nv}", f"-DNP={npar}", f"-DNA={na}", f"-DNW={nw}", f"-DX0SCALE={xs}", f"-DPSCALE={ps}",
           f"-DTEND={tend}", os.path.join(REPO, "oracle", "oracle_model.cpp"),
           os.path.join(REPO, "tests", "sanitize", "oracle_driver.cpp"), "-o", str(exe), f"-L{OMP_LIB}",
           f"-Wl,-rpath,{OMP_LIB}"]
    subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", OMP_NUM_THREADS="2")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-3000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]
    assert "done, failures 0" in out
