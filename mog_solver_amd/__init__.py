"""MI355X-native Mixture-of-Grids (MOG) force step.

Host-side mirror of tbmo/mog-solver's solver interface over hand-written
gfx950 HIP kernels (``lib/libmog.so``, C ABI in ``include/mog.h``).
"""
from .config import SolverConfig, load_config
from .transforms import RandomSE3Sampler, compute_transforms, pack_transforms

__all__ = [
    "SolverConfig",
    "load_config",
    "RandomSE3Sampler",
    "compute_transforms",
    "pack_transforms",
    "MogSolver",
    "Simulation",
]


def __getattr__(name):
    if name in ("MogSolver", "Simulation"):
        from . import solver

        return getattr(solver, name)
    raise AttributeError(name)
