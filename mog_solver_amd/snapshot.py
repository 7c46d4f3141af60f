"""Snapshot save/load of a MOG simulation state.

The reference has a ``record`` key and double-buffered PBOs (config.yaml:15,
main.py:188, 202-204) that are never used; it keeps no state on disk.  A
snapshot here is one ``.npz`` file (NumPy format, no pickled objects):

* ``pos``, ``vel``: (n, 4) float32 particle state in the reference layout
* ``rot``, ``off``: (k, 12) / (k, 4) float32 packed grid transforms
* ``step``: int64 step counter
* ``config``: the solver configuration as a JSON string

``load_snapshot`` reads it back with ``allow_pickle=False``.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

FORMAT_VERSION = 1


def save_snapshot(path, pos, vel, rot, off, step, config: dict):
    """Write one snapshot; device tensors are copied to the host."""
    def host(a):
        if hasattr(a, "detach"):
            a = a.detach().cpu().numpy()
        return np.ascontiguousarray(a, dtype=np.float32)

    cfg = {k: (v if isinstance(v, (int, float, str, bool)) or v is None else str(v)) for k, v in config.items()}
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, pos=host(pos), vel=host(vel), rot=host(rot), off=host(off),
             step=np.int64(step), config=np.array(json.dumps(cfg, sort_keys=True)),
             format_version=np.int64(FORMAT_VERSION))


def load_snapshot(path) -> dict:
    """Read a snapshot into a dict of host arrays (``config`` as a dict)."""
    with np.load(path, allow_pickle=False) as z:
        version = int(z["format_version"]) if "format_version" in z else 0
        if version != FORMAT_VERSION:
            raise ValueError(f"unsupported snapshot format {version}")
        out = {k: z[k] for k in ("pos", "vel", "rot", "off")}
        out["step"] = int(z["step"])
        out["config"] = json.loads(str(z["config"]))
    n = out["pos"].shape[0]
    if out["pos"].shape != (n, 4) or out["vel"].shape != (n, 4):
        raise ValueError("pos/vel must be (n, 4)")
    if out["rot"].shape[1:] != (12,) or out["off"].shape[1:] != (4,) or out["rot"].shape[0] != out["off"].shape[0]:
        raise ValueError("rot/off must be (k, 12) and (k, 4)")
    return out
