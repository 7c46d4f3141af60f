"""Seeded synthetic particle workloads (host NumPy, float32 state).

State layout is the reference's (main.py:191-194): ``pos (n,4) float32 =
(x, y, z, mass)``, ``vel (n,4) float32 = (vx, vy, vz, 0)``, positions in the
periodic cube ``[0, W)``.

* ``uniform_cube``   -- U[0,W)^3, v = 0, unit masses (65,536 particles,
  k=16, g=16 is the small CPU-runnable case).
* ``two_disks``      -- two exponential disks (scale radius 8, sech^2 scale
  height 0.5), centres 40 apart, relative velocity 5 along the separation,
  disk normals 30 degrees apart, circular speeds from the enclosed mass with
  the MOG coupling ``G_eff = G * h^3 / 2`` (density is mass per cell and the
  Poisson factor is -2*pi*G, greens.glsl:40, complex.glsl:23).
* ``plummer_spheres`` -- Plummer spheres at uniform-random centres, v = 0.
* ``plummer_ensemble`` -- 64 uneven simulations (n log-uniform in
  [1e4, 5e5], n_grids in {32, 64, 128, 240}, grid_size in {16, 32}), each a
  Plummer sphere of scale radius 5, for the batched ensemble.
* ``two_galaxies_reference`` -- the reference's key-T state, seeded.
The Perlin-fBm sampler of the reference's key R/1/2/3 runs on the device
(``solver.init_perlin``).
"""
from __future__ import annotations

import numpy as np


def _pack(xyz, vxyz, mass, W):
    n = xyz.shape[0]
    pos = np.zeros((n, 4), dtype=np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    pos[:, :3] = np.mod(xyz, W).astype(np.float32)
    # float32 rounding of a value just below W can give W itself.
    pos[:, :3] = np.where(pos[:, :3] >= np.float32(W), np.float32(0), pos[:, :3])
    pos[:, 3] = mass
    vel[:, :3] = vxyz
    return pos, vel


def uniform_cube(n=65536, world_size=100.0, mass=1.0, seed=1):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, world_size, (n, 3))
    return _pack(xyz, np.zeros((n, 3)), mass, world_size)


def effective_G(G, world_size, grid_size):
    """Newtonian coupling of the MOG force: G * h^3 / 2, h = W / g."""
    h = world_size / grid_size
    return G * h**3 / 2.0


def _disk(rng, n, scale_r, scale_z, normal, center, v_bulk, G_eff, mass):
    R = rng.gamma(2.0, scale_r, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    u = rng.uniform(1e-12, 1 - 1e-12, n)
    z = scale_z * np.arctanh(2 * u - 1)
    local = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    nz = np.asarray(normal, dtype=np.float64)
    nz = nz / np.linalg.norm(nz)
    ref = np.array([1.0, 0.0, 0.0]) if abs(nz[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    ex = np.cross(ref, nz)
    ex /= np.linalg.norm(ex)
    ey = np.cross(nz, ex)
    basis = np.stack([ex, ey, nz], axis=0)
    xyz = local @ basis
    r = np.linalg.norm(xyz, axis=1)
    order = np.argsort(r)
    enclosed = np.empty(n)
    enclosed[order] = np.arange(1, n + 1) * mass
    vc = np.sqrt(G_eff * enclosed / np.maximum(r, 1e-3))
    tang = np.cross(nz[None, :], xyz)
    tang /= np.maximum(np.linalg.norm(tang, axis=1, keepdims=True), 1e-12)
    v = vc[:, None] * tang + np.asarray(v_bulk)[None, :]
    return xyz + np.asarray(center)[None, :], v


def two_disks(n=1_000_000, world_size=100.0, G=100.0, grid_size=16, mass=1.0, seed=2,
              scale_r=8.0, scale_z=0.5, separation=40.0, rel_speed=5.0, incline_deg=30.0):
    """Two colliding exponential disks of n/2 particles each."""
    rng = np.random.default_rng(seed)
    W = world_size
    G_eff = effective_G(G, W, grid_size)
    half = n // 2
    c = W / 2.0
    inc = np.radians(incline_deg)
    a, va = _disk(rng, half, scale_r, scale_z, [0.0, 0.0, 1.0], [c - separation / 2, c, c],
                  [rel_speed / 2, 0, 0], G_eff, mass)
    b, vb = _disk(rng, n - half, scale_r, scale_z, [0.0, np.sin(inc), np.cos(inc)],
                  [c + separation / 2, c, c], [-rel_speed / 2, 0, 0], G_eff, mass)
    return _pack(np.concatenate([a, b]), np.concatenate([va, vb]), mass, W)


def plummer_spheres(n=4_000_000, n_spheres=64, scale=1.5, world_size=100.0, mass=1.0, seed=3):
    """Plummer spheres (radius a / sqrt(u^(-2/3) - 1)) at U(cube) centres."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, world_size, (n_spheres, 3))
    which = rng.integers(0, n_spheres, n)
    u = rng.uniform(1e-6, 1 - 1e-6, n)
    r = scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = np.minimum(r, world_size / 2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    xyz = centres[which] + r[:, None] * d
    return _pack(xyz, np.zeros((n, 3)), mass, world_size)


def plummer_sphere(n, scale=5.0, world_size=100.0, mass=1.0, rng=None, centre=None):
    """One Plummer sphere (radius a / sqrt(u^(-2/3) - 1), isotropic
    directions) about ``centre`` (default the box centre), v = 0."""
    rng = np.random.default_rng() if rng is None else rng
    c = np.full(3, world_size / 2.0) if centre is None else np.asarray(centre, dtype=np.float64)
    u = rng.uniform(1e-6, 1 - 1e-6, n)
    r = np.minimum(scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0), world_size / 2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _pack(c + r[:, None] * d, np.zeros((n, 3)), mass, world_size)


def plummer_ensemble(n_sims=64, n_min=1e4, n_max=5e5, grids=(32, 64, 128, 240), sizes=(16, 32),
                     scale=5.0, world_size=100.0, G=100.0, dt=1e-5, damping=0.999, seed=4):
    """Uneven independent simulations: n = 10^U(log10 n_min, log10 n_max),
    n_grids uniform over ``grids``, grid_size uniform over ``sizes``, each a
    Plummer sphere of scale radius ``scale``.  Returns a list of
    ``(config dict, pos, vel)``."""
    rng = np.random.default_rng(seed)
    ns = np.round(10.0 ** rng.uniform(np.log10(n_min), np.log10(n_max), n_sims)).astype(np.int64)
    ks = rng.choice(np.asarray(grids), n_sims)
    gs = rng.choice(np.asarray(sizes), n_sims)
    out = []
    for i in range(n_sims):
        cfg = dict(n_grids=int(ks[i]), grid_size=int(gs[i]), world_size=world_size, G=G, dt=dt,
                   damping=damping, solver="fft", num_particles=int(ns[i]))
        pos, vel = plummer_sphere(int(ns[i]), scale, world_size, rng=rng)
        out.append((cfg, pos, vel))
    return out


def _spinning_blob(rs, count, world_size, centre, bulk, spin_axis):
    """One blob of ``two_galaxies_reference``: radius ~ Exp(0.08 W), isotropic
    direction (uniform azimuth, uniform cos polar), tangential speed
    0.3 sqrt(r/W + 0.01) about ``spin_axis`` plus a bulk velocity."""
    radius = world_size * rs.exponential(scale=0.08, size=count)
    azimuth = rs.uniform(0, 2 * np.pi, count)
    polar = np.arccos(rs.uniform(-1, 1, count))
    offset = radius[:, None] * np.column_stack(
        (np.sin(polar) * np.cos(azimuth), np.sin(polar) * np.sin(azimuth), np.cos(polar)))
    axis = np.asarray(spin_axis, dtype=np.float64)
    axis /= np.linalg.norm(axis)
    tangent = np.cross(axis, offset)
    tangent /= np.linalg.norm(tangent, axis=1, keepdims=True) + 1e-10
    speed = 0.3 * np.sqrt(radius / world_size + 0.01)
    return centre + offset, speed[:, None] * tangent + np.asarray(bulk)


def two_galaxies_reference(n, world_size, particle_mass=1.0, seed=None):
    """The state of the reference's key T (``Simulation.init_particles2``,
    main.py:312-357): two spinning blobs of n//2 particles at W/2 -+ 0.2 W
    along x with bulk velocities +-(0.3, 0.1, 0) and spin axes along
    (0.3, 0.7, 0.5) and (-0.5, 0.2, 0.8).  The draws follow the reference's
    legacy-RNG order (per blob: radii, azimuths, cos polar) from
    ``RandomState(seed)``; the reference draws from the unseeded global
    state.  Positions are not wrapped into the box (as in the reference)."""
    rs = np.random.RandomState(seed)
    half = n // 2
    mid = world_size / 2.0
    blobs = [
        _spinning_blob(rs, half, world_size, np.array([mid - 0.2 * world_size, mid, mid]),
                       (0.3, 0.1, 0.0), (0.3, 0.7, 0.5)),
        _spinning_blob(rs, half, world_size, np.array([mid + 0.2 * world_size, mid, mid]),
                       (-0.3, -0.1, 0.0), (-0.5, 0.2, 0.8)),
    ]
    pos = np.zeros((n, 4), dtype=np.float32)
    vel = np.zeros((n, 4), dtype=np.float32)
    for i, (xyz, v) in enumerate(blobs):
        pos[i * half:(i + 1) * half, :3] = xyz
        pos[i * half:(i + 1) * half, 3] = particle_mass
        vel[i * half:(i + 1) * half, :3] = v
    return pos, vel
