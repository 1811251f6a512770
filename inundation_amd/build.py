"""Build the HIP backend library in-tree: ``inundation_amd/libinundation_hip.so``.

``python -m inundation_amd.build`` compiles ``csrc/inundation_hip.hip`` for gfx950
with hipcc. FP semantics flags:

- ``-ffp-contract=off``: no implicit multiply-add fusion, so the FP32
  arithmetic is exactly the operations written (the compensated two-sum and
  bit-level parity with the FP32 oracle depend on it);
- no fast-math (correctly rounded division and sqrt are hipcc's defaults);
- ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc" / "inundation_hip.hip"
LIB = PKG / "libinundation_hip.so"
ARCH = os.environ.get("INUNDATION_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose: bool = False) -> Path:
    """Compile the library; always rebuilds. Returns its path."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I",
        str(PKG.parent / "include"),
        "-o",
        str(tmp),
        str(SRC),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
