"""bench.py's command line on the CPU: the one-GPU refusal and --tune."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_parse_tune():
    assert bench.parse_tune(["deposit_wg_target=2048", " spectral_mfma = 1"]) == {
        "deposit_wg_target": 2048, "spectral_mfma": 1}
    assert bench.parse_tune([]) == {}
    for bad in (["deposit_wg_target"], ["deposit_wg_target=abc"], ["=3"], ["x=1.5"]):
        with pytest.raises(SystemExit):
            bench.parse_tune(bad)
