"""Host-side logic: config keys and coercions, transform sampler, C ABI
surface of the built library (no GPU calls beyond validation paths)."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from mog_solver_amd import _abi
from mog_solver_amd.config import FALLBACK, SolverConfig, load_config
from mog_solver_amd.transforms import RandomSE3Sampler, compute_transforms, pack_transforms
from oracle import mog_oracle as O

ROOT = Path(__file__).resolve().parents[1]


# --- config (main.py:115-126, 174-189, 667) ---------------------------------


def test_yaml_string_numbers_are_coerced(tmp_path):
    p = tmp_path / "config.yaml"
    p.write_text("n_grids: 240\ngrid_size: 16\nworld_size: 100.0\nnum_particles: 1e6\nG: 100.0\n"
                 "dt: 1e-5\nsolver: fft\ndamping: 0.999\n")
    raw = load_config(p)
    assert raw["num_particles"] == "1e6" and raw["dt"] == "1e-5"  # PyYAML 1.1 behaviour
    c = SolverConfig.from_dict(raw)
    assert c.num_particles == 1_000_000 and c.dt == 1e-5 and c.n_grids == 240
    assert c.voxel_size == 100.0 / 16 and not c.use_convolution


def test_defaults_and_fallback(tmp_path):
    assert load_config(tmp_path / "missing.yaml") == FALLBACK
    c = SolverConfig.from_dict({"grid_size": 8, "world_size": 10, "G": 1, "dt": 0.1})
    assert c.n_grids == 16 and c.damping == 1.0 and c.solver == "convolution" and c.use_convolution
    assert c.compat == "reference" and c.precision == "fp32" and c.timescale == 1.0


@pytest.mark.parametrize("bad", [{"solver": "direct"}, {"compat": "x"}, {"precision": "fp8"}])
def test_invalid_options(bad):
    d = {"grid_size": 8, "world_size": 10, "G": 1, "dt": 0.1, **bad}
    with pytest.raises(ValueError):
        SolverConfig.from_dict(d)


def test_required_keys():
    with pytest.raises(KeyError):
        SolverConfig.from_dict({"world_size": 10, "G": 1, "dt": 0.1})


# --- transforms --------------------------------------------------------------


@pytest.mark.parametrize("k,g,W,seed", [(240, 16, 100.0, 0), (16, 16, 100.0, 0), (9, 32, 2000.0, 5)])
def test_sampler_matches_oracle_bit_exact(k, g, W, seed):
    off, rot = compute_transforms(k, g, W, seed=seed)
    orot, ooff = O.se3_transforms(k, g, W, seed)
    assert np.array_equal(rot, orot) and np.array_equal(off, ooff)
    mats, offs = RandomSE3Sampler(g, W, k, seed).generate()
    r2, o2 = pack_transforms(mats, offs)
    assert np.array_equal(r2, rot) and np.array_equal(o2, off)


# --- C ABI -------------------------------------------------------------------


def header_functions():
    text = (ROOT / "include" / "mog.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mog(?:2d)?_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    assert _abi.LIB_PATH.exists(), "build() first"
    declared = header_functions()
    assert set(declared) == set(_abi.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mog(?:2d)?_\w+)", out))
    missing = set(declared) - exported
    assert not missing, missing
    lib = _abi.load()
    for name in declared:
        assert hasattr(lib, name)
    version = re.search(r"#define MOG_ABI_VERSION (\d+)", (ROOT / "include" / "mog.h").read_text())
    assert lib.mog_abi_version() == int(version.group(1)) == _abi.ABI_VERSION


def test_config_struct_layout():
    assert ctypes.sizeof(_abi.MogConfig) == 64
    assert _abi.MogConfig.world_size.offset == 8 and _abi.MogConfig.compat.offset == 48


def _cfg(**kw):
    base = dict(n_grids=16, grid_size=16, world_size=100.0, G=100.0, dt=1e-5, damping=1.0, timescale=1.0,
                compat=0, precision=0, solver=0)
    base.update(kw)
    return _abi.MogConfig(**base)


@pytest.mark.parametrize("kw,code", [
    (dict(grid_size=12), _abi.MOG_EINVAL),
    (dict(grid_size=128), _abi.MOG_EINVAL),
    (dict(grid_size=1), _abi.MOG_EINVAL),
    (dict(n_grids=0), _abi.MOG_EINVAL),
    (dict(n_grids=70000), _abi.MOG_EINVAL),
    (dict(world_size=-1.0), _abi.MOG_EINVAL),
    (dict(world_size=float("nan")), _abi.MOG_EINVAL),
    (dict(dt=float("inf")), _abi.MOG_EINVAL),
    (dict(compat=5), _abi.MOG_EINVAL),
    (dict(precision=9), _abi.MOG_EINVAL),
    (dict(solver=3), _abi.MOG_EINVAL),
    (dict(precision=1, grid_size=8), _abi.MOG_EUNSUPPORTED),
    (dict(precision=2, grid_size=4), _abi.MOG_EUNSUPPORTED),
])
def test_create_rejects_invalid_config(kw, code):
    lib = _abi.load()
    h = ctypes.c_void_p()
    rc = lib.mog_create(ctypes.byref(_cfg(**kw)), ctypes.byref(h))
    assert rc == code
    assert h.value is None
    assert lib.mog_last_error(None)


def test_null_arguments():
    lib = _abi.load()
    assert lib.mog_create(None, None) == _abi.MOG_EINVAL
    h = ctypes.c_void_p()
    assert lib.mog_create(None, ctypes.byref(h)) == _abi.MOG_EINVAL
    assert lib.mog_step(None, None, None, 10, None) == _abi.MOG_EINVAL
    assert lib.mog_deposit(None, None, 10, None, None) == _abi.MOG_EINVAL
    assert lib.mog_set_transforms(None, None, None) == _abi.MOG_EINVAL
    assert lib.mog_grids_ptr(None) is None
    lib.mog_destroy(None)
    assert b"NULL" in lib.mog_last_error(None)
