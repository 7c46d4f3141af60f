"""Host check that the fast NGP cell path is exact (no GPU).

Compiles tests/native/cellmath_check.cpp against the kernels' own
mog_solver_amd/csrc/cellmath.h with g++ (-ffp-contract=off) and runs it on
uniform and adversarial on-face positions for W in {100, 2000, 10, 1, 37.3},
g in {2..64} and the admission bounds |c| <= W/2 (3D step) and 1.5 W (planar
path), and checks that the shifted-field form of the fast cell
(the byte offsets the LDS deposit and the gather use) gives the same cell
wherever its range condition admits it.  A run with the margin removed must
report mismatches, which shows the check can fail.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("cellmath") / "cellmath_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(ROOT / "mog_solver_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "cellmath_check.cpp"), "-o", str(exe)], check=True)
    return exe


def test_fast_cell_path_is_exact(checker):
    r = subprocess.run([str(checker), "10000000"], capture_output=True, text=True)
    bad, rate, total, shifted = r.stdout.split()
    assert r.returncode == 0 and int(bad) == 0, r.stderr
    # both admission bounds; a third of the samples lie outside W/2 on purpose
    assert int(total) > 19_000_000 and float(rate) > 0.25
    assert int(shifted) > 15_000_000


def test_check_detects_missing_margin(checker):
    r = subprocess.run([str(checker), "2000000", "0"], capture_output=True, text=True)
    assert r.returncode == 1 and int(r.stdout.split()[0]) > 0
