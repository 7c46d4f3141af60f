"""Python solver interface over the HIP C ABI.

``MogSolver`` owns one ``mog_ctx`` (device grids, tables, transforms) and
runs the phases or whole steps on device tensors.  ``Simulation`` is the
headless counterpart of the reference's ``Simulation`` class
(main.py:101-672): same configuration keys, attribute names
(``cfg``, ``n_grids``, ``grid_size``, ``world_size``, ``voxel_size``,
``num_particles``, ``G``, ``dt``, ``timescale``, ``offsets``, ``rotations``,
``use_convolution``, ``pos_buf``, ``vel_buf``) and step methods
(``step``, ``step_fft``, ``clear_mass``, ``compute_transforms``,
``load_config``), without the window, renderer and key handling.

Device tensors are torch tensors on a HIP device; torch supplies memory and
the current stream only, every computation runs in ``libmog.so``.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import _abi
from .config import SolverConfig, load_config
from .transforms import compute_transforms


def _torch():
    import torch

    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def init_perlin(pos, vel, world_size, noise_scale=4.0, density_contrast=2.0, seed=12345,
                base_mass=1.0, spawn_buffer=0.1, spawn_shape=1, stream=None):
    """Perlin-fBm initial conditions into device tensors pos, vel (n,4)
    (shaders/init_particles.glsl through mog_init_perlin)."""
    _check_particles(pos, "pos")
    _check_particles(vel, "vel")
    lib = _abi.load()
    rc = lib.mog_init_perlin(_ptr(pos), _ptr(vel), pos.shape[0], float(world_size), float(noise_scale),
                             float(density_contrast), int(seed) & 0xFFFFFFFF, float(base_mass),
                             float(spawn_buffer), int(spawn_shape), _stream_ptr(stream))
    _abi.check(lib, None, rc)


def mog_config(cfg: SolverConfig) -> _abi.MogConfig:
    """The C ``mog_config`` of a coerced solver config."""
    return _abi.MogConfig(
        n_grids=cfg.n_grids,
        grid_size=cfg.grid_size,
        world_size=cfg.world_size,
        G=cfg.G,
        dt=cfg.dt,
        damping=cfg.damping,
        timescale=cfg.timescale,
        compat=_abi.COMPAT[cfg.compat],
        precision=_abi.PRECISION[cfg.precision],
        solver=_abi.SOLVER[cfg.solver],
        reserved=0,
    )


def _check_particles(t, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 4 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous (n,4) float32 tensor, got {tuple(t.shape)} {t.dtype}")


class MogSolver:
    """One MOG force-step context (``mog_create`` ... ``mog_destroy``)."""

    def __init__(self, cfg: SolverConfig, rot=None, off=None, seed: int = 0):
        self.lib = _abi.load()
        self.cfg = cfg
        c = mog_config(cfg)
        h = ctypes.c_void_p()
        rc = self.lib.mog_create(ctypes.byref(c), ctypes.byref(h))
        _abi.check(self.lib, None, rc)
        self._ctx = h
        if rot is None or off is None:
            off, rot = compute_transforms(cfg.n_grids, cfg.grid_size, cfg.world_size, seed=seed)
        self.set_transforms(rot, off)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.mog_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        _abi.check(self.lib, self._ctx, rc)

    # -- configuration ------------------------------------------------------
    def set_transforms(self, rot, off):
        """Upload packed (k,12) rotations and (k,4) offsets (host arrays)."""
        rot = np.ascontiguousarray(rot, dtype=np.float32)
        off = np.ascontiguousarray(off, dtype=np.float32)
        k = self.cfg.n_grids
        if rot.shape != (k, 12) or off.shape != (k, 4):
            raise ValueError(f"expected rot (k,12) and off (k,4) with k={k}, got {rot.shape}, {off.shape}")
        self.rot = rot
        self.off = off
        self._check(self.lib.mog_set_transforms(self._ctx, rot.ctypes.data, off.ctypes.data))

    def set_integrator(self, dt=None, damping=None, timescale=None):
        """Change dt, damping or timescale between steps."""
        c = self.cfg
        c.dt = c.dt if dt is None else float(dt)
        c.damping = c.damping if damping is None else float(damping)
        c.timescale = c.timescale if timescale is None else float(timescale)
        self._check(self.lib.mog_set_integrator(self._ctx, c.dt, c.damping, c.timescale))

    # -- tuning -------------------------------------------------------------
    def get_tuning(self) -> dict:
        t = _abi.MogTuning()
        self._check(self.lib.mog_get_tuning(self._ctx, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **fields):
        """Update tuning fields (see ``struct mog_tuning``); others keep their value."""
        cur = self.get_tuning()
        unknown = set(fields) - set(cur)
        if unknown:
            raise KeyError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(fields)
        t = _abi.MogTuning(**{k: int(v) for k, v in cur.items()})
        self._check(self.lib.mog_set_tuning(self._ctx, ctypes.byref(t)))

    @property
    def cells(self):
        g = self.cfg.grid_size
        return self.cfg.n_grids * g * g * g

    # -- whole step ---------------------------------------------------------
    def step(self, pos, vel, stream=None):
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        if pos.shape[0] != vel.shape[0]:
            raise ValueError("pos and vel must have the same length")
        self._check(self.lib.mog_step(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0], _stream_ptr(stream)))

    def step_noalloc(self, pos, vel, stream=None):
        """One step that allocates no device memory (``mog_step_noalloc``:
        n <= 2^20, bit-identical to ``step``)."""
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        if pos.shape[0] != vel.shape[0]:
            raise ValueError("pos and vel must have the same length")
        self._check(self.lib.mog_step_noalloc(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0], _stream_ptr(stream)))

    def run(self, pos, vel, nsteps, use_graph=True, stream=None):
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        if pos.shape[0] != vel.shape[0]:
            raise ValueError("pos and vel must have the same length")
        self._check(
            self.lib.mog_run(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0], int(nsteps), int(bool(use_graph)),
                             _stream_ptr(stream))
        )

    def step_profiled(self, pos, vel, stream=None):
        """One step; returns the HIP-event durations (ms) of its phases
        {clear, deposit, solve, gather_update} (synchronises the stream).
        "clear" holds no kernel: the previous step's gather cleared the grids
        this one deposits into."""
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        ms = (ctypes.c_float * 4)()
        self._check(self.lib.mog_step_profiled(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0],
                                               ctypes.cast(ms, ctypes.c_void_p), _stream_ptr(stream)))
        return dict(zip(("clear", "deposit", "solve", "gather_update"), list(ms)))

    # -- phases -------------------------------------------------------------
    def deposit(self, pos, out=None, stream=None):
        """int64 milli-mass grids (k, g, g, g) [grid][z][y][x]."""
        torch = _torch()
        _check_particles(pos, "pos")
        g = self.cfg.grid_size
        if out is None:
            out = torch.empty((self.cfg.n_grids, g, g, g), dtype=torch.int64, device=pos.device)
        self._check(self.lib.mog_deposit(self._ctx, _ptr(pos), pos.shape[0], _ptr(out), _stream_ptr(stream)))
        return out

    def solve(self, grids, out=None, stream=None):
        """float32 force grids (k, g, g, g, 4) = (Fx, Fy, Fz, |F|)."""
        torch = _torch()
        g = self.cfg.grid_size
        if grids.dtype != torch.int64 or grids.numel() != self.cells or not grids.is_contiguous():
            raise ValueError("grids must be a contiguous int64 tensor of k*g^3 elements")
        if out is None:
            out = torch.empty((self.cfg.n_grids, g, g, g, 4), dtype=torch.float32, device=grids.device)
        self._check(self.lib.mog_solve(self._ctx, _ptr(grids), _ptr(out), _stream_ptr(stream)))
        return out

    def gather(self, pos, force, out=None, stream=None):
        """avgForce (n,4) float32 (w = 0) from force grids."""
        torch = _torch()
        _check_particles(pos, "pos")
        if force.dtype != torch.float32 or force.numel() != 4 * self.cells or not force.is_contiguous():
            raise ValueError("force must be a contiguous float32 tensor of k*g^3*4 elements")
        if out is None:
            out = torch.empty_like(pos)
        self._check(self.lib.mog_gather(self._ctx, _ptr(pos), pos.shape[0], _ptr(force), _ptr(out),
                                        _stream_ptr(stream)))
        return out

    def update(self, pos, vel, avg_force, stream=None):
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        _check_particles(avg_force, "avg_force")
        self._check(self.lib.mog_update(self._ctx, _ptr(pos), _ptr(vel), _ptr(avg_force), pos.shape[0],
                                        _stream_ptr(stream)))

    def forces(self, pos, stream=None):
        """avgForce of the whole pipeline (deposit, solve, gather) at ``pos``."""
        grids = self.deposit(pos, stream=stream)
        force = self.solve(grids, stream=stream)
        return self.gather(pos, force, stream=stream)

    def diagnostics(self, pos, vel, stream=None) -> dict:
        """FP64 momentum (3,), kinetic energy, grid potential energy (mean
        over the k grids of 1/2 sum rho phi, the potential whose gradient
        the force grids hold) and total mass of a state (``mog_diagnostics``;
        synchronises the stream)."""
        torch = _torch()
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        if pos.shape[0] != vel.shape[0]:
            raise ValueError("pos and vel must have the same length")
        out = torch.zeros(8, dtype=torch.float64, device=pos.device)
        self._check(self.lib.mog_diagnostics(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0], _ptr(out),
                                             _stream_ptr(stream)))
        o = out.cpu().numpy()
        return {"momentum": o[0:3].copy(), "kinetic": float(o[3]), "potential": float(o[4]),
                "mass": float(o[5])}


class Simulation:
    """Headless MOG simulation with the reference's configuration and names.

    ``config`` is a dict of config.yaml keys or a path to such a file (the
    reference's fallback dict when the path does not exist).  Particle state
    is set with ``init_particles(pos, vel)`` from host arrays or device
    tensors of shape (n,4).
    """

    def __init__(self, config=None, seed: int = 0, device=None, transforms=None):
        self.cfg = config if isinstance(config, dict) else self.load_config(config)
        sc = SolverConfig.from_dict(self.cfg)
        self.solver_cfg = sc
        self.timescale = sc.timescale
        self.paused = False
        self.n_grids = sc.n_grids
        self.grid_size = sc.grid_size
        self.world_size = sc.world_size
        self.voxel_size = self.world_size / self.grid_size
        self.num_particles = sc.num_particles
        self.G = sc.G
        self.dt = sc.dt
        self.seed = seed
        self.step_count = 0
        if transforms is None:
            self.offsets, self.rotations = self.compute_transforms()
        else:
            rot, off = transforms
            self.rotations = np.ascontiguousarray(rot, dtype=np.float32)
            self.offsets = np.ascontiguousarray(off, dtype=np.float32)
        self.use_convolution = sc.use_convolution
        self.device = device
        self.solver = MogSolver(sc, rot=self.rotations, off=self.offsets)
        self.pos_buf = None
        self.vel_buf = None

    @staticmethod
    def load_config(path=None):
        return load_config(path)

    def compute_transforms(self):
        """Packed (offsets (k,4), rotations (k,12)) from the seeded sampler."""
        return compute_transforms(self.n_grids, self.grid_size, self.world_size, seed=self.seed)

    def init_particles(self, pos, vel=None):
        torch = _torch()
        dev = self.device or torch.device("cuda", torch.cuda.current_device())
        pos_t = torch.as_tensor(np.asarray(pos, dtype=np.float32) if not isinstance(pos, torch.Tensor) else pos)
        pos_t = pos_t.to(device=dev, dtype=torch.float32).contiguous()
        if vel is None:
            vel_t = torch.zeros_like(pos_t)
        else:
            vel_t = torch.as_tensor(np.asarray(vel, dtype=np.float32) if not isinstance(vel, torch.Tensor) else vel)
            vel_t = vel_t.to(device=dev, dtype=torch.float32).contiguous()
        self.pos_buf, self.vel_buf = pos_t, vel_t
        self.num_particles = pos_t.shape[0]

    # Presets of the reference's keys 1/2/3 (main.py:722-731).
    PRESETS = {1: (2.0, 1.5), 2: (4.0, 2.5), 3: (8.0, 4.0)}

    def _alloc_state(self):
        torch = _torch()
        dev = self.device or torch.device("cuda", torch.cuda.current_device())
        n = self.num_particles
        if self.pos_buf is None or self.pos_buf.shape[0] != n:
            self.pos_buf = torch.empty((n, 4), dtype=torch.float32, device=dev)
            self.vel_buf = torch.empty((n, 4), dtype=torch.float32, device=dev)

    def init_particles_gpu(self, noise_scale=4.0, density_contrast=2.0, seed=None, spawn_buffer=None):
        """Perlin-fBm initial conditions on the device (main.py:283-310)."""
        if seed is None:
            seed = int(np.random.randint(0, 2**31))
        if spawn_buffer is None:
            spawn_buffer = self.cfg.get("spawn_buffer", 0.1)
        self._alloc_state()
        init_perlin(self.pos_buf, self.vel_buf, self.world_size, noise_scale, density_contrast, seed,
                    float(self.cfg.get("particle_mass", 1.0)), float(spawn_buffer), 1)

    def init_particles_default(self):
        """The reference's start-up and key R state (main.py:280-281)."""
        self.init_particles_gpu(noise_scale=4.0, density_contrast=2.0, seed=12345)

    def preset(self, number, seed=None):
        """Keys 1/2/3: clustering presets."""
        ns, dc = self.PRESETS[int(number)]
        self.init_particles_gpu(noise_scale=ns, density_contrast=dc, seed=seed)

    def init_particles2(self, seed=None):
        """Key T: two galaxies (main.py:312-357)."""
        from .workloads import two_galaxies_reference

        pos, vel = two_galaxies_reference(self.num_particles, self.world_size,
                                          float(self.cfg.get("particle_mass", 1.0)), seed)
        self.init_particles(pos, vel)

    def set_timescale(self, timescale):
        """deltaTime = dt * timescale (keys [ ] \\ of the reference)."""
        self.timescale = float(timescale)
        self.solver.set_integrator(timescale=self.timescale)

    def clear_mass(self):
        """Kept for interface parity: the step clears its grids itself."""

    def step_fft(self):
        if self.pos_buf is None:
            raise RuntimeError("init_particles() first")
        if self.solver.cfg.timescale != self.timescale:  # attribute assigned directly
            self.solver.set_integrator(timescale=self.timescale)
        self.solver.step(self.pos_buf, self.vel_buf)

    def step(self):
        """The reference always runs the FFT solver (main.py:671-672)."""
        if not self.paused:
            self.step_fft()
            self.step_count += 1

    def run(self, nsteps, use_graph=True):
        """``nsteps`` steps.  With ``record`` set in the config, a snapshot is
        written every ``record_every`` steps (default 10) to ``record_dir``
        (default ``snapshots``)."""
        if self.pos_buf is None:
            raise RuntimeError("init_particles() first")
        if self.paused:
            return
        if self.solver.cfg.timescale != self.timescale:
            self.solver.set_integrator(timescale=self.timescale)
        every = int(self.cfg.get("record_every", 10)) if self.cfg.get("record") else 0
        done = 0
        while done < nsteps:
            chunk = nsteps - done if every <= 0 else min(nsteps - done, every - self.step_count % every)
            self.solver.run(self.pos_buf, self.vel_buf, chunk, use_graph=use_graph)
            done += chunk
            self.step_count += chunk
            if every > 0 and self.step_count % every == 0:
                self.save(Path(self.cfg.get("record_dir", "snapshots")) / f"step_{self.step_count:08d}.npz")

    def save(self, path):
        """Write the current state as a snapshot (``mog_solver_amd.snapshot``)."""
        from .snapshot import save_snapshot

        save_snapshot(path, self.pos_buf, self.vel_buf, self.rotations, self.offsets, self.step_count,
                      self.cfg)

    @classmethod
    def from_snapshot(cls, path, device=None):
        """A simulation restored from a snapshot: configuration, transforms,
        particle state and step counter."""
        from .snapshot import load_snapshot

        snap = load_snapshot(path)
        sim = cls(snap["config"], device=device, transforms=(snap["rot"], snap["off"]))
        sim.init_particles(snap["pos"], snap["vel"])
        sim.step_count = snap["step"]
        return sim
