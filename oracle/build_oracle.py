"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/oracle_model.cpp once per right-hand-side model into
oracle/_build/oracle_<model>.so with ROCm's clang++ and the same
floating-point contraction rule as the hipRTC programs (-ffp-contract=on,
hardware FMA).  OpenMP parallelises over ensemble members.

Used by tests/oracle.py, __graft_entry__.build()/smoke() and bench.py's
cpu_baseline leg.  The library never loads these objects.

The reference's own C++ core is not built: it needs spdlog/fmt through Bazel
(network) and an OpenCL device, neither of which exists here.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ORACLE = Path(__file__).resolve().parent
REPO = ORACLE.parent
BUILD = ORACLE / "_build"
MODELS = REPO / "clode_amd" / "models"
CLANG = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "llvm" / "bin" / "clang++"
OMP_LIB = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "llvm" / "lib"

DEFAULT_MODELS = ["van_der_pol", "lorenz", "lactotroph", "lactotroph_noise", "chay_keizer",
                  "ornstein_uhlenbeck", "sine", "linear_drag", "harmonic", "decay", "two_tone"]


# deliberately wrong checkers ("mutants"), built only by the tests that must
# tell them apart from the oracle: variant name -> compiler defines
MUTANTS = {"noshift": ["-DCLODE_ORACLE_MUTANT_NO_TIME_SHIFT"],
           # the operation-counting build (flop_count.hpp): precision 9 counts
           "count": ["-DCLODE_ORACLE_COUNTING"]}


def oracle_path(model: str, variant: str = "") -> Path:
    return BUILD / f"oracle_{model}{'_' + variant if variant else ''}.so"


def build_model(model: str, rhs: Path | None = None, force: bool = False, variant: str = "") -> Path:
    rhs = Path(rhs) if rhs is not None else MODELS / f"{model}.cl"
    out = oracle_path(model, variant)
    srcs = [ORACLE / "oracle_model.cpp", ORACLE / "clode_oracle.hpp", ORACLE / "oracle_ensemble.hpp",
            ORACLE / "flop_count.hpp",
            REPO / "clode_amd" / "csrc" / "device" / "cm_math.h",
            REPO / "clode_amd" / "csrc" / "device" / "clode_builtins.h", rhs]
    if out.exists() and not force and all(out.stat().st_mtime >= s.stat().st_mtime for s in srcs):
        return out
    BUILD.mkdir(exist_ok=True)
    cmd = [str(CLANG), "-O2", "-std=c++17", "-ffp-contract=on", "-mfma", "-fopenmp", "-fPIC", "-shared",
           "-Wno-unused-parameter", "-Wno-unused-variable",
           f'-DCLODE_ORACLE_RHS="{rhs}"', *MUTANTS.get(variant, []), str(ORACLE / "oracle_model.cpp"), "-o", str(out),
           f"-L{OMP_LIB}", f"-Wl,-rpath,{OMP_LIB}"]
    subprocess.run(cmd, check=True)
    return out


# models of the benchmark workloads, whose operation-counting builds generate
# and check clode_amd/perfmodel.py (gen_perfmodel.py)
COUNT_MODELS = ["van_der_pol", "lorenz", "lactotroph", "lactotroph_noise", "chay_keizer", "ornstein_uhlenbeck"]


def build_all(force: bool = False) -> list[Path]:
    return ([build_model(m, force=force) for m in DEFAULT_MODELS] +
            [build_model(m, force=force, variant="count") for m in COUNT_MODELS])


if __name__ == "__main__":
    for p in build_all(force="--force" in sys.argv):
        print(p)
