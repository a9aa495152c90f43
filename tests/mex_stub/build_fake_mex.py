"""Builds the three MATLAB gateways (clode_amd/matlab/clode_mex.cpp) against
the stand-in MEX runtime of this folder, as shared libraries the tests drive
through ctypes: tests/mex_stub/_build/libfake_{clODEmex,clODEfeaturesmex,
clODEtrajectorymex}.so, linked to clode_amd/libclode_hip.so."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
OUT = HERE / "_build"
NAMES = {0: "clODEmex", 1: "clODEfeaturesmex", 2: "clODEtrajectorymex"}


def lib_path(kind: int) -> Path:
    return OUT / f"libfake_{NAMES[kind]}.so"


def build(kind: int) -> Path:
    OUT.mkdir(exist_ok=True)
    out = lib_path(kind)
    cmd = ["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-Wall", f"-DCLODE_MEX_KIND={kind}", f"-I{HERE}",
           f"-I{REPO / 'include'}", str(REPO / "clode_amd" / "matlab" / "clode_mex.cpp"), str(HERE / "fake_mex.cpp"),
           f"-L{REPO / 'clode_amd'}", "-lclode_hip", "-Wl,-rpath,$ORIGIN/../../../clode_amd", "-o", str(out)]
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


def build_all() -> list[Path]:
    return [build(k) for k in NAMES]


if __name__ == "__main__":
    build_all()
    sys.exit(0)
