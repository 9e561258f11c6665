"""Build recipe of the oracle library (test infrastructure only).

Compiles oracle/ckm1d_oracle.c with gcc into oracle/build/libckm1d_oracle.so
(OpenMP over systems).  The reference itself has no C/C++ source for this
path (CUDA kernels behind CuPy/NVRTC and Numba-CUDA, plus a NumPy/SciPy/Numba
CPU script), so there is no oracle/_ref build.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "ckm1d_oracle.c"
LIB = HERE / "build" / "libckm1d_oracle.so"


def build(verbose: bool = False) -> Path:
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
           "-o", str(tmp), str(SRC), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    tmp.replace(LIB)
    return LIB


def ensure_built() -> Path:
    if not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
        build()
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
