"""Compile the FP32 oracle: ``oracle/libcpu_ref.so`` (TEST INFRASTRUCTURE).

g++ with OpenMP, no FP contraction and no fast math, so every float operation
in ``cpu_ref.cpp`` is the single IEEE operation written.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_ref.cpp"
LIB = HERE / "libcpu_ref.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        os.environ.get("CXX", "g++"),
        "-O2",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-fopenmp",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-fno-math-errno",
        "-o",
        str(tmp),
        str(SRC),
    ]
    res = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
