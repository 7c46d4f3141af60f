"""Build recipe of the HIP library ``libmog.so`` (gfx950).

``build()`` compiles every ``csrc/*.hip`` translation unit with hipcc (in
parallel, one process per unit, objects in a temporary directory) and links
them into one shared library inside the package
(``mog_solver_amd/lib/libmog.so``), so the built file travels with the
source tree.  Flags:

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
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libmog.so"
ARCH = os.environ.get("MOG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "deposit.hip", "spectral.hip", "spectral_mfma.hip", "gather.hip", "order.hip", "init.hip",
           "closure.hip", "batch.hip", "mog2d.hip", "energy.hip", "noalloc.hip"]

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]
# kept for callers that print or reuse the flag list
HIPCC_FLAGS = [*COMMON_FLAGS, "-shared"]

_NOISE = "is not a recognized feature for this target"


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if _NOISE not in l)
    return res, err


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile the library; raises CalledProcessError on failure."""
    LIBDIR.mkdir(exist_ok=True)
    sources = list(SOURCES)
    jobs = jobs or max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with tempfile.TemporaryDirectory(prefix="mog_build_") as tmpdir:
        objs = [str(Path(tmpdir) / (Path(s).stem + ".o")) for s in sources]
        cmds = [[hipcc(), *COMMON_FLAGS, "-c", str(CSRC / s), "-o", o] for s, o in zip(sources, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda c: (c, *_run(c, verbose)), cmds))
        for cmd, res, err in results:
            if res.returncode != 0:
                sys.stderr.write(err + "\n")
                raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
            if verbose and err.strip():
                sys.stderr.write(err + "\n")
        tmp = LIB.with_suffix(".so.tmp")
        link = [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *objs, "-o", str(tmp)]
        res, err = _run(link, verbose)
        if res.returncode != 0:
            sys.stderr.write(err + "\n")
            raise subprocess.CalledProcessError(res.returncode, link, res.stdout, res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
