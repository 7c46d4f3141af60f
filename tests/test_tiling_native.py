"""Host check of the workgroup-to-tile map the deposit, gather and closure
feature kernels launch with (no GPU).

Compiles tests/native/tiling_check.cpp against the kernels' own
mog_solver_amd/csrc/tiling.h with g++ and runs it for every tile count up to
20000 and a few large ones: each map must be a bijection whose tiles, per
XCD (workgroups L with L % 8 equal), form consecutive ranges in order of L.
"""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_xcd_tile_is_a_contiguous_bijection(tmp_path):
    exe = tmp_path / "tiling_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "mog_solver_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "tiling_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "20000"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.split() == ["0"], r.stdout + r.stderr
