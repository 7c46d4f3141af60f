"""ctypes binding of ``libmog.so`` (the C ABI declared in ``include/mog.h``).

The library is built in-tree by ``mog_solver_amd/build.py`` into
``mog_solver_amd/lib/libmog.so``.  Loading fails loudly when it is missing:
there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmog.so"
ABI_VERSION = 2  # MOG_ABI_VERSION of include/mog.h this binding follows

MOG_OK = 0
MOG_EINVAL = -1
MOG_EHIP = -2
MOG_ENOTRANSFORMS = -3
MOG_EUNSUPPORTED = -4
MOG_ENOMEM = -5

COMPAT = {"reference": 0, "physical": 1}
PRECISION = {"fp32": 0, "bf16x3": 1, "fp16": 2}
SOLVER = {"fft": 0, "convolution": 1}

# Every symbol include/mog.h declares (checked by tests against the .so).
EXPORTED = (
    "mog_abi_version",
    "mog_create",
    "mog_destroy",
    "mog_set_transforms",
    "mog_step",
    "mog_run",
    "mog_deposit",
    "mog_solve",
    "mog_gather",
    "mog_update",
    "mog_step_profiled",
    "mog_step_noalloc",
    "mog_set_integrator",
    "mog_init_perlin",
    "mog_diagnostics",
    "mog_closure_features",
    "mog_mlp_param_count",
    "mog_mlp_forward",
    "mog_mlp_train_grad",
    "mog_mlp_adam",
    "mog_direct_sum",
    "mog_get_tuning",
    "mog_set_tuning",
    "mog_grids_ptr",
    "mog_force_ptr",
    "mog_batch_create",
    "mog_batch_destroy",
    "mog_batch_set_transforms",
    "mog_batch_step",
    "mog_batch_run",
    "mog_batch_set_tuning",
    "mog_batch_get_tuning",
    "mog2d_create",
    "mog2d_destroy",
    "mog2d_forces",
    "mog2d_deposit",
    "mog2d_step",
    "mog_last_error",
)


class MogConfig(ctypes.Structure):
    """Mirror of ``struct mog_config`` in include/mog.h."""

    _fields_ = [
        ("n_grids", ctypes.c_int32),
        ("grid_size", ctypes.c_int32),
        ("world_size", ctypes.c_double),
        ("G", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("timescale", ctypes.c_double),
        ("compat", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("solver", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class MogTuning(ctypes.Structure):
    """Mirror of ``struct mog_tuning`` in include/mog.h."""

    _fields_ = [
        ("deposit_threads", ctypes.c_int32),
        ("deposit_grids_per_wg", ctypes.c_int32),
        ("deposit_wg_target", ctypes.c_int32),
        ("resort_every", ctypes.c_int32),
        ("sorted_deposit", ctypes.c_int32),
        ("cell_cache", ctypes.c_int32),
        ("spectral_mfma", ctypes.c_int32),
        ("deposit_prepass", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 8),
    ]

    FIELDS = ("deposit_threads", "deposit_grids_per_wg", "deposit_wg_target", "resort_every",
              "sorted_deposit", "cell_cache", "spectral_mfma", "deposit_prepass")

    def as_dict(self):
        return {f: int(getattr(self, f)) for f in self.FIELDS}


class MogError(RuntimeError):
    """A negative status returned by the library."""

    def __init__(self, code: int, message: str):
        super().__init__(f"mog error {code}: {message}")
        self.code = code


_lib = None


def load(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Load (once) and return the library with argument types declared."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"HIP library {p} is not built; run `python -c \"import __graft_entry__ as g; g.build()\"` "
            "or `python mog_solver_amd/build.py`"
        )
    lib = ctypes.CDLL(str(p))
    vp = ctypes.c_void_p
    i64 = ctypes.c_int64
    i32 = ctypes.c_int32
    lib.mog_abi_version.restype = ctypes.c_int
    lib.mog_abi_version.argtypes = []
    if lib.mog_abi_version() != ABI_VERSION:
        raise ImportError(f"{p} has ABI version {lib.mog_abi_version()}, this binding needs {ABI_VERSION}; "
                          "rebuild it")
    lib.mog_create.restype = ctypes.c_int
    lib.mog_create.argtypes = [ctypes.POINTER(MogConfig), ctypes.POINTER(vp)]
    lib.mog_destroy.restype = None
    lib.mog_destroy.argtypes = [vp]
    lib.mog_set_transforms.restype = ctypes.c_int
    lib.mog_set_transforms.argtypes = [vp, vp, vp]
    lib.mog_step.restype = ctypes.c_int
    lib.mog_step.argtypes = [vp, vp, vp, i64, vp]
    lib.mog_run.restype = ctypes.c_int
    lib.mog_run.argtypes = [vp, vp, vp, i64, i32, i32, vp]
    lib.mog_deposit.restype = ctypes.c_int
    lib.mog_deposit.argtypes = [vp, vp, i64, vp, vp]
    lib.mog_solve.restype = ctypes.c_int
    lib.mog_solve.argtypes = [vp, vp, vp, vp]
    lib.mog_gather.restype = ctypes.c_int
    lib.mog_gather.argtypes = [vp, vp, i64, vp, vp, vp]
    lib.mog_update.restype = ctypes.c_int
    lib.mog_update.argtypes = [vp, vp, vp, vp, i64, vp]
    lib.mog_step_noalloc.restype = ctypes.c_int
    lib.mog_step_noalloc.argtypes = [vp, vp, vp, i64, vp]
    lib.mog_step_profiled.restype = ctypes.c_int
    lib.mog_step_profiled.argtypes = [vp, vp, vp, i64, vp, vp]
    lib.mog_set_integrator.restype = ctypes.c_int
    lib.mog_set_integrator.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.mog_init_perlin.restype = ctypes.c_int
    lib.mog_init_perlin.argtypes = [vp, vp, i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint32, ctypes.c_double, ctypes.c_double, i32, vp]
    d = ctypes.c_double
    lib.mog_diagnostics.restype = ctypes.c_int
    lib.mog_diagnostics.argtypes = [vp, vp, vp, i64, vp, vp]
    lib.mog_closure_features.restype = ctypes.c_int
    lib.mog_closure_features.argtypes = [vp, vp, i64, d, vp, vp, vp]
    lib.mog_mlp_param_count.restype = ctypes.c_int
    lib.mog_mlp_param_count.argtypes = []
    lib.mog_mlp_forward.restype = ctypes.c_int
    lib.mog_mlp_forward.argtypes = [vp, vp, vp, i64, vp, i32, vp]
    lib.mog_mlp_train_grad.restype = ctypes.c_int
    lib.mog_mlp_train_grad.argtypes = [vp, vp, vp, vp, i64, d, vp, vp, i32, vp]
    lib.mog_mlp_adam.restype = ctypes.c_int
    lib.mog_mlp_adam.argtypes = [vp, vp, vp, vp, d, d, d, d, i32, vp]
    lib.mog_direct_sum.restype = ctypes.c_int
    lib.mog_direct_sum.argtypes = [vp, i64, vp, i32, d, d, d, vp, vp]
    lib.mog_get_tuning.restype = ctypes.c_int
    lib.mog_get_tuning.argtypes = [vp, ctypes.POINTER(MogTuning)]
    lib.mog_set_tuning.restype = ctypes.c_int
    lib.mog_set_tuning.argtypes = [vp, ctypes.POINTER(MogTuning)]
    lib.mog_grids_ptr.restype = vp
    lib.mog_grids_ptr.argtypes = [vp]
    lib.mog_force_ptr.restype = vp
    lib.mog_force_ptr.argtypes = [vp]
    lib.mog_batch_create.restype = ctypes.c_int
    lib.mog_batch_create.argtypes = [ctypes.POINTER(MogConfig), vp, i32, ctypes.POINTER(vp)]
    lib.mog_batch_destroy.restype = None
    lib.mog_batch_destroy.argtypes = [vp]
    lib.mog_batch_set_transforms.restype = ctypes.c_int
    lib.mog_batch_set_transforms.argtypes = [vp, i32, vp, vp]
    lib.mog_batch_step.restype = ctypes.c_int
    lib.mog_batch_step.argtypes = [vp, vp, vp, vp]
    lib.mog_batch_run.restype = ctypes.c_int
    lib.mog_batch_run.argtypes = [vp, vp, vp, i32, i32, vp]
    lib.mog_batch_set_tuning.restype = ctypes.c_int
    lib.mog_batch_set_tuning.argtypes = [vp, i64, i32]
    lib.mog_batch_get_tuning.restype = ctypes.c_int
    lib.mog_batch_get_tuning.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.mog2d_create.restype = ctypes.c_int
    lib.mog2d_create.argtypes = [i32, i32, d, d, vp, vp, ctypes.POINTER(vp)]
    lib.mog2d_destroy.restype = None
    lib.mog2d_destroy.argtypes = [vp]
    lib.mog2d_forces.restype = ctypes.c_int
    lib.mog2d_forces.argtypes = [vp, vp, i64, vp, vp]
    lib.mog2d_deposit.restype = ctypes.c_int
    lib.mog2d_deposit.argtypes = [vp, vp, i64, vp, vp]
    lib.mog2d_step.restype = ctypes.c_int
    lib.mog2d_step.argtypes = [vp, vp, vp, i64, d, d, vp]
    lib.mog_last_error.restype = ctypes.c_char_p
    lib.mog_last_error.argtypes = [vp]
    if path is None:
        _lib = lib
    return lib


def check(lib, ctx, rc: int) -> None:
    """Raise MogError for a negative status, with the library's message."""
    if rc != MOG_OK:
        msg = lib.mog_last_error(ctx)
        raise MogError(rc, msg.decode() if msg else "")
