"""Build recipe for the CPU oracle (test infrastructure, see pqb_oracle.cpp).

    python -m oracle.build        ->  oracle/_build/libpqb_oracle.so

Compiled with -ffp-contract=off so that double and long double arithmetic is
unfused, as in the reference's release flags (-O3 -m64 -mavx2, no -mfma).
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
LIB = OUT / "libpqb_oracle.so"
SRC = HERE / "pqb_oracle.cpp"

CXX = os.environ.get("CXX", "g++")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-Wall"]


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [CXX, *FLAGS, str(SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
