"""``config.yaml`` handling with the reference's keys and coercions.

Keys and defaults follow ``Simulation.__init__`` (main.py:115-126) and
``step_fft`` (main.py:667): ``n_grids`` (default 16), ``grid_size``,
``world_size``, ``num_particles``, ``G``, ``dt`` (required), ``solver``
(default ``"convolution"``), ``damping`` (default 1.0), plus
``particle_mass`` / ``spawn_buffer`` / ``record`` which the force step does
not read.  The file is read with ``yaml.safe_load`` and, when absent,
replaced by the fallback dict of main.py:180-189.

PyYAML (YAML 1.1) returns ``3e6`` and ``1e-5`` as *strings*; the reference
coerces ``num_particles`` with ``int(float(...))`` (main.py:119) but uses
``dt`` raw (main.py:121), which would fail at ``dt * timescale``.  Here every
numeric key is coerced with ``float(...)``, so such files work.

Two keys are this build's own: ``compat`` (``reference`` keeps the x<-z
force aliasing of the reference at g in {4, 16, 64}; ``physical`` uses the
true x force) and ``precision`` (``fp32``; ``bf16x3`` and ``fp16`` name the
split-precision DFT variants).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import yaml

FALLBACK = {
    "n_grids": 16,
    "grid_size": 16,
    "world_size": 100.0,
    "num_particles": 100000,
    "G": 1.0,
    "dt": 0.00001,
    "particle_mass": 1.0,
    "solver": "fft",
    "record": True,
}


def load_config(path: str | Path | None = None) -> dict:
    """``yaml.safe_load`` of ``path``; the reference's fallback dict if the
    file does not exist (main.py:174-189)."""
    if path is not None:
        p = Path(path)
        if p.exists():
            with open(p) as f:
                data = yaml.safe_load(f)
            return dict(data or {})
    return dict(FALLBACK)


def _num(v: Any) -> float:
    return float(v)


@dataclass
class SolverConfig:
    """Coerced solver parameters (attribute names of main.py:115-123)."""

    n_grids: int
    grid_size: int
    world_size: float
    num_particles: int
    G: float
    dt: float
    damping: float = 1.0
    solver: str = "convolution"
    timescale: float = 1.0
    compat: str = "reference"
    precision: str = "fp32"
    particle_mass: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def voxel_size(self) -> float:
        return self.world_size / self.grid_size

    @property
    def use_convolution(self) -> bool:
        return self.solver == "convolution"

    @staticmethod
    def from_dict(cfg: dict) -> "SolverConfig":
        known = {
            "n_grids", "grid_size", "world_size", "num_particles", "G", "dt", "damping",
            "solver", "timescale", "compat", "precision", "particle_mass",
        }
        solver = str(cfg.get("solver", "convolution"))
        if solver not in ("fft", "convolution"):
            raise ValueError(f"solver must be 'fft' or 'convolution', got {solver!r}")
        compat = str(cfg.get("compat", "reference"))
        if compat not in ("reference", "physical"):
            raise ValueError(f"compat must be 'reference' or 'physical', got {compat!r}")
        precision = str(cfg.get("precision", "fp32"))
        if precision not in ("fp32", "bf16x3", "fp16"):
            raise ValueError(f"precision must be fp32, bf16x3 or fp16, got {precision!r}")
        return SolverConfig(
            n_grids=int(cfg.get("n_grids", 16)),
            grid_size=int(cfg["grid_size"]),
            world_size=_num(cfg["world_size"]),
            num_particles=int(_num(cfg.get("num_particles", 0))),
            G=_num(cfg["G"]),
            dt=_num(cfg["dt"]),
            damping=_num(cfg.get("damping", 1.0)),
            solver=solver,
            timescale=_num(cfg.get("timescale", 1.0)),
            compat=compat,
            precision=precision,
            particle_mass=_num(cfg.get("particle_mass", 1.0)),
            extra={k: v for k, v in cfg.items() if k not in known},
        )
