"""CPU twin of the Perlin-fBm initial-condition sampler -- TEST INFRASTRUCTURE.

Restates shaders/init_particles.glsl:20-287 (3D branch) in NumPy, vectorised
over particles, to check the HIP sampler (mog_init_perlin):

* integer hash chain (``hash``/``hash2``/``hash3``, init_particles.glsl:20-35)
  in exact uint32 arithmetic -> bit-exact;
* ``rand`` = float32(h) / float32(4294967295.0) (init_particles.glsl:37);
* Perlin gradients, fade, mix and 4-octave fBm (init_particles.glsl:53-133)
  in float32 with NumPy's float32 transcendentals;
* rejection sampling in a sphere (spawnShape 1) or cube (0), up to 64
  attempts, uniform fallback (init_particles.glsl:221-285).

Pinning: the hash is checked against values computed by an independent
scalar Python implementation of the same integer recipe and against the
avalanche property; the sampler's acceptance rate is checked against the
density field's mean, and its output distribution against the density
field (particles concentrate where fbm is high).  Transcendental rounding
differs between NumPy and the device library by an ulp, so a candidate
within an ulp of its threshold may be decided differently: parity with the
HIP kernel is exact for velocities and for all but a tiny fraction of
positions.
"""
from __future__ import annotations

import numpy as np

U32 = np.uint32
F32 = np.float32


def ihash(x):
    x = np.asarray(x, dtype=U32)
    with np.errstate(over="ignore"):
        x = x ^ (x >> U32(16))
        x = (x * U32(0x85EBCA6B)).astype(U32)
        x = x ^ (x >> U32(13))
        x = (x * U32(0xC2B2AE35)).astype(U32)
        x = x ^ (x >> U32(16))
    return x


def ihash2(x, y):
    x = np.asarray(x, dtype=U32)
    with np.errstate(over="ignore"):
        return ihash(x ^ (ihash(y) + U32(0x9E3779B9) + (x << U32(6)) + (x >> U32(2))))


def ihash3(x, y, z):
    with np.errstate(over="ignore"):
        return ihash(ihash2(x, y) ^ (ihash(z) + U32(0x9E3779B9)))


def urand(s):
    return (ihash(s).astype(F32) / F32(4294967295.0)).astype(F32)


def urand3(s):
    h1 = ihash(s)
    h2 = ihash(h1)
    return urand(s), urand(h1), urand(h2)


def _u32(x):
    return np.asarray(x).astype(np.int64).astype(U32) if np.asarray(x).dtype.kind == "i" else np.asarray(x, U32)


def gradient(ix, iy, iz, s):
    s = U32(s)
    with np.errstate(over="ignore"):
        h = ihash3(_u32(ix) + s, _u32(iy) + s * U32(7), _u32(iz) + s * U32(13))
    theta = ((h & U32(0xFFFF)).astype(F32) * F32(6.28318530718)) / F32(65535.0)
    phi = np.arccos(((F32(2.0) * (h >> U32(16)).astype(F32)) / F32(65535.0)) - F32(1.0)).astype(F32)
    sp = np.sin(phi).astype(F32)
    return (sp * np.cos(theta)).astype(F32), (sp * np.sin(theta)).astype(F32), np.cos(phi).astype(F32)


def fade(t):
    t = t.astype(F32)
    return (t * t * t * (t * (t * F32(6.0) - F32(15.0)) + F32(10.0))).astype(F32)


def mix(x, y, a):
    return (x * (F32(1.0) - a) + y * a).astype(F32)


def perlin(px, py, pz, s):
    fx0, fy0, fz0 = np.floor(px), np.floor(py), np.floor(pz)
    ix, iy, iz = fx0.astype(np.int64), fy0.astype(np.int64), fz0.astype(np.int64)
    fx, fy, fz = (px - fx0).astype(F32), (py - fy0).astype(F32), (pz - fz0).astype(F32)
    ux, uy, uz = fade(fx), fade(fy), fade(fz)
    n = []
    for c in range(8):
        dx, dy, dz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        gx, gy, gz = gradient(ix + dx, iy + dy, iz + dz, s)
        ox, oy, oz = fx - F32(dx), fy - F32(dy), fz - F32(dz)
        n.append(((gx * ox + gy * oy) + gz * oz).astype(F32))
    nx00, nx10 = mix(n[0], n[1], ux), mix(n[2], n[3], ux)
    nx01, nx11 = mix(n[4], n[5], ux), mix(n[6], n[7], ux)
    return mix(mix(nx00, nx10, uy), mix(nx01, nx11, uy), uz)


def fbm(px, py, pz, s):
    value = np.zeros_like(px, dtype=F32)
    amp, freq = F32(0.5), F32(1.0)
    for i in range(4):
        value = (value + amp * perlin((px * freq).astype(F32), (py * freq).astype(F32),
                                      (pz * freq).astype(F32), U32(s) + U32(i) * U32(1000))).astype(F32)
        amp = F32(amp * F32(0.5))
        freq = F32(freq * F32(2.0))
    return value


def init_perlin(n, world_size, noise_scale=4.0, density_contrast=2.0, seed=12345, base_mass=1.0,
                spawn_buffer=0.1, spawn_shape=1):
    """(pos (n,4), vel (n,4), attempts (n,)) float32 initial conditions."""
    W = F32(world_size)
    ns, dc = F32(noise_scale), F32(density_contrast)
    with np.errstate(over="ignore"):
        ps = (U32(seed) + np.arange(n, dtype=U32) * U32(12345)).astype(U32)
    minb = F32(F32(spawn_buffer) * W)
    maxb = F32((F32(1.0) - F32(spawn_buffer)) * W)
    rng = F32(maxb - minb)
    center = F32(W * F32(0.5))
    radius = F32(rng * F32(0.5))
    pos = np.zeros((n, 3), F32)
    done = np.zeros(n, bool)
    attempts = np.full(n, 64, np.int64)
    for a in range(64):
        idx = np.nonzero(~done)[0]
        if idx.size == 0:
            break
        base = (ps[idx] + U32(a) * U32(1000)).astype(U32)
        if spawn_shape == 1:
            sp = np.zeros((idx.size, 3), F32)
            pending = np.ones(idx.size, bool)
            for sa in range(16):
                j = np.nonzero(pending)[0]
                if j.size == 0:
                    break
                r = urand3((base[j] + U32(sa) * U32(100)).astype(U32))
                for d in range(3):
                    sp[j, d] = r[d] * F32(2.0) - F32(1.0)
                d2 = (sp[j, 0] * sp[j, 0] + sp[j, 1] * sp[j, 1]) + sp[j, 2] * sp[j, 2]
                pending[j] = d2 > F32(1.0)
            cand = (center + sp * radius).astype(F32)
        else:
            r = urand3(base)
            cand = np.stack([(ri * rng + minb).astype(F32) for ri in r], axis=1)
        nc = [((cand[:, d] * ns) / W).astype(F32) for d in range(3)]
        dens = (fbm(nc[0], nc[1], nc[2], seed) + F32(0.5)).astype(F32)
        dens = np.power(np.clip(dens, F32(0.0), F32(1.0)), dc).astype(F32)
        thr = urand((base + U32(500)).astype(U32))
        acc = thr < dens
        pos[idx[acc]] = cand[acc]
        attempts[idx[acc]] = a
        done[idx[acc]] = True
    rest = np.nonzero(~done)[0]
    if rest.size:
        if spawn_shape == 1:
            r = (np.power(urand(ps[rest]), F32(1.0 / 3.0)) * radius).astype(F32)
            th = (urand(ihash(ps[rest])) * F32(6.28318530718)).astype(F32)
            ph = np.arccos(F32(2.0) * urand(ihash(ihash(ps[rest]))) - F32(1.0)).astype(F32)
            pos[rest, 0] = center + np.sin(ph) * np.cos(th) * r
            pos[rest, 1] = center + np.sin(ph) * np.sin(th) * r
            pos[rest, 2] = center + np.cos(ph) * r
        else:
            r = urand3(ps[rest])
            for d in range(3):
                pos[rest, d] = r[d] * rng + minb
    v = urand3((ps + U32(999)).astype(U32))
    out_p = np.zeros((n, 4), F32)
    out_v = np.zeros((n, 4), F32)
    out_p[:, :3] = pos
    out_p[:, 3] = F32(base_mass)
    for d in range(3):
        out_v[:, d] = ((v[d] - F32(0.5)) * F32(1.0)) * W
    return out_p, out_v, attempts
