"""Autotuner of the MOG step's launch and algorithm choices.

Times whole steps (``MogSolver.run`` with graph replay, as the step is run
in production; the periodic rebuild of the visiting order happens between
replays and is included) on copies of the caller's state for candidate
``mog_tuning`` settings, and keeps the fastest.  Each candidate is timed as
the best of three runs; the three fastest of a stage are then timed again
over runs four times longer before the stage's winner is taken, so that one
lucky measurement cannot pick a slower tiling.  The search is coordinate
descent over the tuning fields, starting from the library's defaults:

1. deposit tiling: threads per workgroup x grids per workgroup x workgroups
2. steps between rebuilds of the Morton visiting order (0 = storage order)
3. deposit in visiting order (wave-aggregated adds) and the cell cache
4. matrix-core or vector DFT for FP32

Stages 1-3 do not change results: the step is bit-identical across the
deposit tiling, visiting-order and cell-cache choices.  Stage 4 does: the
matrix-core FP32 DFT rounds differently from the vector FFTs, so the forces
it gives differ from the default path's by FP32 rounding (both stay within
the solve's tolerance against the FP64 oracle, a few 1e-7 relative RMS).
Winners are cached in a JSON file keyed by device, n, k, g and precision.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

DEFAULT_CACHE = Path(os.environ.get("XDG_CACHE_HOME", str(Path.home() / ".cache"))) / "mog_solver_amd" / "tuning.json"


def _key(solver, n: int) -> str:
    import torch

    name = torch.cuda.get_device_name(torch.cuda.current_device())
    c = solver.cfg
    return f"{name}|n={n}|k={c.n_grids}|g={c.grid_size}|prec={c.precision}|compat={c.compat}"


def time_steps(solver, pos, vel, steps: int = 8, reps: int = 3) -> float:
    """Milliseconds per step over ``steps`` graph-replayed steps on copies of
    pos/vel (best of ``reps`` runs, after one warm-up step that captures)."""
    import torch

    p = pos.clone()
    v = vel.clone()
    solver.run(p, v, 1, use_graph=True)  # warm: order, buffers, graph capture
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        start.record()
        solver.run(p, v, steps, use_graph=True)
        stop.record()
        torch.cuda.synchronize()
        best = min(best, start.elapsed_time(stop) / steps)
    return best


def candidate_stages(g: int):
    grid_bytes = 4 * g**3
    tiles = []
    for threads in (256, 512, 1024):
        for G in (1, 2, 4, 8):
            if G > 1 and G * grid_bytes > 160 * 1024:
                continue
            for wg in (128, 256, 512, 1024, 2048, 4096):
                tiles.append(dict(deposit_threads=threads, deposit_grids_per_wg=G, deposit_wg_target=wg))
    return [
        tiles,
        [dict(resort_every=r) for r in (0, 8, 16, 32)],
        [dict(sorted_deposit=0, cell_cache=0), dict(sorted_deposit=1, cell_cache=0),
         dict(sorted_deposit=1, cell_cache=1)],
        [dict(spectral_mfma=1), dict(spectral_mfma=0)],
    ]


def autotune(solver, pos, vel, steps: int = 8, cache_path: os.PathLike | None = DEFAULT_CACHE,
             use_cache: bool = True, verbose: bool = False) -> dict:
    """Pick and apply the fastest tuning for this state's size; returns
    ``{"tuning": ..., "ms_per_step": ..., "evaluated": ...}``."""
    n = int(pos.shape[0])
    key = _key(solver, n)
    cache = {}
    if cache_path is not None and use_cache:
        try:
            cache = json.loads(Path(cache_path).read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            solver.set_tuning(**cache[key]["tuning"])
            return dict(cache[key], cached=True)
    best = solver.get_tuning()
    evaluated = 0
    g = solver.cfg.grid_size
    for stage in candidate_stages(g):
        timed = []  # (ms, tuning), the incumbent included
        for cand in [{}] + list(stage):
            trial = {**best, **cand}
            if cand and trial == best:
                continue
            try:
                solver.set_tuning(**trial)
            except Exception:  # rejected combination
                continue
            ms = time_steps(solver, pos, vel, steps)
            evaluated += 1
            if verbose:
                print(f"[autotune] {cand} -> {ms:.4f} ms/step", flush=True)
            timed.append((ms, trial))
        # confirm the three fastest over longer runs
        finalists = sorted(timed, key=lambda t: t[0])[:3]
        confirmed = []
        for _, trial in finalists:
            solver.set_tuning(**trial)
            confirmed.append((time_steps(solver, pos, vel, 4 * steps), trial))
        best_ms, best = min(confirmed, key=lambda t: t[0])
        solver.set_tuning(**best)
    result = {"tuning": best, "ms_per_step": best_ms, "evaluated": evaluated}
    if cache_path is not None:
        cache[key] = result
        try:
            Path(cache_path).parent.mkdir(parents=True, exist_ok=True)
            Path(cache_path).write_text(json.dumps(cache, indent=1, sort_keys=True))
        except OSError:
            pass
    return dict(result, cached=False)
