"""Build recipe of the HIP library ``libmog.so`` (gfx950).

``build()`` compiles every ``csrc/*.hip`` translation unit with hipcc into
one shared library inside the package (``mog_solver_amd/lib/libmog.so``),
so the built file travels with the source tree.  Flags:

* ``-ffp-contract=off``: the cell-index arithmetic must round every product
  and sum separately to reproduce the reference's float32 sequence bit for
  bit; code that wants FMA calls ``fmaf`` explicitly.
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE float32 division in the
  ``mod``/``int(l / h)`` steps of the cell index.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libmog.so"
ARCH = os.environ.get("MOG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "deposit.hip", "spectral.hip", "spectral_mfma.hip", "gather.hip", "order.hip", "init.hip", "closure.hip", "batch.hip", "mog2d.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def build(verbose: bool = False) -> Path:
    """Compile the library; raises CalledProcessError on failure."""
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        sys.stderr.write(err + "\n")
        raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    if verbose and err.strip():
        sys.stderr.write(err + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
