"""CPU oracle for the Mixture-of-Grids (MOG) force step -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``mog_solver_amd``) runs HIP kernels and fails
loudly when its extension is missing; it never falls back to this file.

It restates the reference pipeline of tbmo/mog-solver in NumPy:

* transforms   -- ``main.py:17-62`` (RandomSE3Sampler) and ``main.py:157-172``
* cell index   -- ``shaders/mass.glsl:38-50`` == ``shaders/update.glsl:67-76``,
                  emulated in float32 with a fixed operation order and no FMA,
                  so integer density grids are bit-identical to the kernels'
* deposit      -- ``shaders/mass.glsl:52-54`` (integer milli-mass atomics)
* solve        -- ``shaders/complex.glsl``, ``fft_batched.glsl``,
                  ``fourier.glsl``, ``greens.glsl:25-61``, ``pack.glsl`` and the
                  buffer ping-pong of ``main.py:564-590``/``638-654``, in FP64
* gather       -- ``shaders/update.glsl:67-81`` in FP64
* update       -- ``shaders/update.glsl:83-89`` in float32 (same op order as
                  the kernel, so given the same averaged force it is bit-exact)

Pinning (what each function is checked against, beyond its own formula) is
listed per function below and exercised in ``tests/test_oracle.py``:

* ``se3_transforms``  -- orthonormality/det, a known quaternion, the GLSL
  column-major ``mat3`` constructor semantics, offsets range, RNG draw order.
  The bit pattern of the NumPy PCG64 stream is pinned only to the installed
  NumPy version (the reference pins none).
* ``cell_indices``    -- identity-transform known answers, a 90-degree
  rotation known answer (pins the transpose), agreement with an FP64
  computation away from cell faces, brute-force per-particle loop.
* ``deposit``         -- per-grid mass conservation, permutation invariance,
  brute-force ``np.add.at``.
* ``solve``           -- closed-form single Fourier mode, brute-force direct
  DFT at small g, radix-2 restatement vs ``numpy.fft``, odd-kernel invariants
  (zero net momentum), the buffer-aliasing behaviour of the reference.
* ``gather``/``update`` -- identity reductions to single-grid PM, zero-force
  closed form, rotated uniform field.
* ``solve_convolution`` -- brute-force pair sum at small g, the lone-mass
  closed form ``G voxel^3 / 2 * m r / |r|^3``, zero net momentum for a
  density confined to less than half the box, and agreement with the FFT
  solver's force at separations of several cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

F32 = np.float32

# ---------------------------------------------------------------------------
# A.1 Transforms  (main.py:17-62, main.py:157-172)
# ---------------------------------------------------------------------------


def quat_to_matrix(q):
    """Rotation matrix of quaternion (w, x, y, z) in float32 arithmetic.

    Follows ``RandomSE3Sampler._quat_to_matrix`` (main.py:37-62): the entries
    are evaluated from float32 quaternion components (the quaternions were
    cast with ``astype(np.float32)`` at main.py:26) and stored in a float32
    array; the reference then returns the transpose's first three rows,
    i.e. ``M.T[:3, :4]``.  This function returns the 3x3 ``M`` itself.
    """
    w, x, y, z = (F32(v) for v in q)
    one, two = F32(1), F32(2)
    m = np.array(
        [
            [one - two * y * y - two * z * z, two * x * y - two * z * w, two * x * z + two * y * w],
            [two * x * y + two * z * w, one - two * x * x - two * z * z, two * y * z - two * x * w],
            [two * x * z - two * y * w, two * y * z + two * x * w, one - two * x * x - two * y * y],
        ],
        dtype=np.float32,
    )
    return m


def se3_transforms(n_grids: int, grid_size: int, world_size: float, seed=0):
    """Per-grid rotations and offsets in the reference's packed GPU layouts.

    Restates ``RandomSE3Sampler.generate`` (main.py:17-30) and
    ``compute_transforms`` (main.py:157-172):

    * ``u = rng.uniform(0, 1, (k, 3))`` then quaternions (Shoemake) in FP64,
      cast to float32;
    * offsets ``rng.uniform(0, voxel, (k, 3))`` drawn *after* ``u`` from the
      same generator, cast to float32;
    * packed rotation ``rot[i, 4*j + m] = M_i[j][m]`` (j, m < 3; slots 3, 7,
      11 are zero), packed offset ``off[i] = (ox, oy, oz, 0)``.

    Returns ``(rot (k,12) float32, off (k,4) float32)``.
    """
    k = int(n_grids)
    voxel = float(world_size) / int(grid_size)
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, (k, 3))
    q0 = np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1])
    q1 = np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1])
    q2 = np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2])
    q3 = np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])
    quats = np.stack([q0, q1, q2, q3], axis=1).astype(np.float32)
    offs = rng.uniform(0, voxel, (k, 3)).astype(np.float32)

    rot = np.zeros((k, 12), dtype=np.float32)
    off = np.zeros((k, 4), dtype=np.float32)
    for i in range(k):
        m = quat_to_matrix(quats[i])
        for j in range(3):
            rot[i, 4 * j : 4 * j + 3] = m[j, :]
        off[i, :3] = offs[i]
    return rot, off


def rotation_matrices(rot):
    """(k,12) packed layout -> (k,3,3) matrices M with rot[4j+m] = M[j][m]."""
    rot = np.asarray(rot, dtype=np.float32)
    return np.stack([rot[:, 0:3], rot[:, 4:7], rot[:, 8:11]], axis=1)


# ---------------------------------------------------------------------------
# Scalar parameters exactly as the shaders receive them
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class Params:
    """Uniform values as float32, the way main.py passes them to the shaders.

    ``voxelSize`` is computed in Python double (main.py:118) and rounded to
    float32 when set as a uniform (main.py:604, 665); ``worldSize`` likewise
    (main.py:605, 664); ``deltaTime = dt * timescale`` (main.py:662).
    """

    k: int
    g: int
    W: np.float32
    half: np.float32
    h: np.float32

    @staticmethod
    def make(n_grids, grid_size, world_size):
        W = F32(world_size)
        return Params(
            k=int(n_grids),
            g=int(grid_size),
            W=W,
            half=F32(W * F32(0.5)),
            h=F32(float(world_size) / int(grid_size)),
        )


# ---------------------------------------------------------------------------
# A.2 Cell index  (mass.glsl:42-50 == update.glsl:71-76 + 44-48)
# ---------------------------------------------------------------------------


def glsl_mod32(x, y):
    """GLSL ``mod(x, y) = x - y * floor(x / y)`` in float32, no FMA."""
    x = np.asarray(x, dtype=np.float32)
    y = F32(y)
    t = np.floor(x / y).astype(np.float32)
    return (x - (y * t).astype(np.float32)).astype(np.float32)


def axis_cell32(l, W, h, g):
    """One axis of the NGP cell: ``int(mod(mod(l,W)+W,W) / h) % g``.

    ``int()`` truncates toward zero (GLSL float->int); the truncated value
    is non-negative for every finite input, so ``% g`` equals ``& (g-1)``.
    Non-finite inputs map to an unspecified in-bounds cell.
    """
    l = glsl_mod32(l, W)
    l = (l + W).astype(np.float32)
    l = glsl_mod32(l, W)
    q = (l / h).astype(np.float32)
    with np.errstate(invalid="ignore"):
        i = np.trunc(q)
        i = np.where(np.isfinite(i), i, 0).astype(np.int64)
    return i & (g - 1)


def centered32(pos, P: Params):
    """``basePos - worldSize * 0.5`` (mass.glsl:42), float32, shape (n,3)."""
    p = np.asarray(pos, dtype=np.float32)[:, :3]
    return (p - P.half).astype(np.float32)


def cell_indices_one(c, rot_i, off_i, P: Params):
    """Flat cell index ``((z*g)+y)*g+x`` of every particle in one grid.

    ``c`` is the centred float32 position (n,3); ``rot_i``/``off_i`` the
    packed 12-/4-float rows of this grid.  The local position is
    ``rot * centered`` with ``rot = mat3(col0, col1, col2)`` built from the
    packed rows (mass.glsl:20-25, 43), i.e. ``r_m = sum_j rot[4j+m] * c_j``,
    evaluated as ``(rot[m]*c0 + rot[4+m]*c1) + rot[8+m]*c2`` with every
    product and sum rounded to float32 (the kernels use the same order and
    no contraction).  Then ``l = (r + W/2) - off`` (mass.glsl:44).
    """
    R = np.asarray(rot_i, dtype=np.float32)
    o = np.asarray(off_i, dtype=np.float32)
    c0, c1, c2 = c[:, 0], c[:, 1], c[:, 2]
    idx = np.zeros(c.shape[0], dtype=np.int64)
    shift = {0: 0, 1: 1, 2: 2}
    lg = int(round(math.log2(P.g)))
    for m in range(3):
        r = (R[m] * c0).astype(np.float32)
        r = (r + (R[4 + m] * c1).astype(np.float32)).astype(np.float32)
        r = (r + (R[8 + m] * c2).astype(np.float32)).astype(np.float32)
        l = (r + P.half).astype(np.float32)
        l = (l - o[m]).astype(np.float32)
        idx |= axis_cell32(l, P.W, P.h, P.g) << (lg * shift[m])
    return idx


def cell_indices(pos, rot, off, P: Params):
    """(k, n) int64 flat cell indices for all grids."""
    c = centered32(pos, P)
    return np.stack([cell_indices_one(c, rot[i], off[i], P) for i in range(P.k)])


# ---------------------------------------------------------------------------
# A.3 Deposit  (mass.glsl:52-54)
# ---------------------------------------------------------------------------


def milli_mass(pos):
    """``uint(mass * 1000.0)`` as int64: float32 product, truncation.

    Negative or NaN masses deposit 0 (the reference's conversion of a
    negative float to uint is undefined).
    """
    m = np.asarray(pos, dtype=np.float32)[:, 3]
    mm = (m * F32(1000.0)).astype(np.float32)
    with np.errstate(invalid="ignore"):
        mm = np.where(mm > 0, mm, F32(0))
    return np.trunc(mm.astype(np.float64)).astype(np.int64)


def _bincount_i64(idx, w, size):
    """Exact int64 weighted bincount (np.bincount weights go through f64)."""
    out = np.zeros(size, dtype=np.int64)
    np.add.at(out, idx, w)
    return out


def deposit(pos, rot, off, P: Params):
    """(k,g,g,g) int64 milli-mass grids, order-independent by construction."""
    c = centered32(pos, P)
    milli = milli_mass(pos)
    g = P.g
    out = np.zeros((P.k, g, g, g), dtype=np.int64)
    uniform = milli.size > 0 and np.all(milli == milli[0])
    for i in range(P.k):
        idx = cell_indices_one(c, rot[i], off[i], P)
        if uniform:
            out[i] = (np.bincount(idx, minlength=g**3).astype(np.int64) * milli[0]).reshape(g, g, g)
        else:
            out[i] = _bincount_i64(idx, milli, g**3).reshape(g, g, g)
    return out


# ---------------------------------------------------------------------------
# A.4 Solve  (complex.glsl, fft_batched.glsl, fourier.glsl, greens.glsl,
#             pack.glsl, main.py:592-655)
# ---------------------------------------------------------------------------


def aliases_x_with_z(g: int) -> bool:
    """True where the reference's x force is overwritten by its z force.

    ``run_fft_3d_batched`` (main.py:564-590) ping-pongs src/dst once per
    dispatch and runs ``log2(g)+1`` dispatches per axis, 3 axes, returning
    the last destination.  With an odd dispatch count the result lands in
    the ``dst`` argument; the x and z inverse transforms (main.py:638-646)
    both pass ``complex_tex_b`` as ``dst``, so z overwrites x and
    ``pack.glsl`` reads (Fz, Fy, Fz).  Odd count <=> log2(g) even.
    """
    stages = int(round(math.log2(g)))
    return (3 * (stages + 1)) % 2 == 1


def wave_numbers(g, W):
    """FP64 ``k_i = s(i) * 2*pi / W`` with ``s(i) = i if i <= g/2 else i-g``
    (greens.glsl:17-21; the Nyquist index maps to +g/2)."""
    s = np.array([i if i <= g // 2 else i - g for i in range(g)], dtype=np.float64)
    return s * (2.0 * np.pi / float(W))


def greens_mask32(g, W, eps=1e-4):
    """``k2 > epsilon`` evaluated as greens.glsl:32-45 does, in float32.

    ``PI = 3.14159265359`` (greens.glsl:14) rounds to float32(pi);
    ``k = float(shifted) * (TWO_PI / L)``; ``k2 = kx*kx + ky*ky + kz*kz``.
    Returned as a (g,g,g) [z][y][x] boolean array.
    """
    two_pi = F32(2.0) * F32(3.14159265359)
    step = F32(two_pi / F32(W))
    s = np.array([i if i <= g // 2 else i - g for i in range(g)], dtype=np.float32)
    kk = (s * step).astype(np.float32)
    kz = kk[:, None, None]
    ky = kk[None, :, None]
    kx = kk[None, None, :]
    k2 = ((kx * kx).astype(np.float32) + (ky * ky).astype(np.float32)).astype(np.float32)
    k2 = (k2 + (kz * kz).astype(np.float32)).astype(np.float32)
    return k2 > F32(eps)


def greens_factor(g, W, G):
    """FP64 Green's function ``-2*pi*G / k^2`` on the (g,g,g) [z][y][x]
    spectrum, zero where the float32 ``k2 > 1e-4`` test fails (DC always).
    ``geoFactor = -2*PI`` (greens.glsl:40)."""
    k = wave_numbers(g, W)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    mask = greens_mask32(g, W)
    with np.errstate(divide="ignore", invalid="ignore"):
        gf = np.where(mask, -2.0 * np.pi * float(G) / np.where(k2 > 0, k2, 1.0), 0.0)
    gf[0, 0, 0] = 0.0  # fourier.glsl:16-18 zeroes DC before greens
    return gf


def _radix2_axis(x, axis, direction, flush):
    """One axis of the reference's radix-2 DIT (fft_batched.glsl:45-132).

    Stage 0 is the bit-reversal permutation, stages 1..log2(g) the
    butterflies ``out[i1] = p + q*w``, ``out[i2] = p - q*w`` with
    ``w = exp(-direction * 2*pi*i * k / stageSize)``.  With ``flush`` the
    twiddle components (fft_batched.glsl:29-32) and the product components
    (fft_batched.glsl:123-126) of magnitude below 1e-6 are set to zero.
    Computed in complex128.
    """
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, -1).copy()
    n = x.shape[-1]
    nb = int(round(math.log2(n)))
    rev = np.array([int(format(i, f"0{nb}b")[::-1], 2) if nb > 0 else 0 for i in range(n)])
    y = np.empty_like(x)
    y[..., rev] = x
    x = y
    for stage in range(1, nb + 1):
        size = 1 << stage
        half = size >> 1
        t = np.arange(n // 2)
        i1 = (t // half) * size + (t % half)
        i2 = i1 + half
        kk = t % half
        phase = -direction * 2.0 * np.pi * kk / size
        c = np.cos(phase)
        s = np.sin(phase)
        c = np.where(kk == 0, 1.0, np.where(kk == size // 2, -1.0, c))
        s = np.where(kk == 0, 0.0, np.where(kk == size // 2, 0.0, s))
        if flush:
            c = np.where(np.abs(c) < 1e-6, 0.0, c)
            s = np.where(np.abs(s) < 1e-6, 0.0, s)
        w = c + 1j * s
        p = x[..., i1]
        q = x[..., i2]
        temp = q * w
        if flush:
            tr = np.where(np.abs(temp.real) < 1e-6, 0.0, temp.real)
            ti = np.where(np.abs(temp.imag) < 1e-6, 0.0, temp.imag)
            temp = tr + 1j * ti
        out = np.empty_like(x)
        out[..., i1] = p + temp
        out[..., i2] = p - temp
        x = out
    return np.moveaxis(x, -1, axis)


def fft3_radix2(x, direction, flush=False):
    """Unnormalised 3D transform over the last three axes, x then y then z
    (main.py:574: axes 0,1,2 = X,Y,Z of the [..., z, y, x] array)."""
    for axis in (-1, -2, -3):
        x = _radix2_axis(x, axis, direction, flush)
    return x


def solve(grids, g, W, G, compat="reference", fft="numpy", flush=False, solver="fft"):
    """Force grids (k,g,g,g,4) float64 = (Fx, Fy, Fz, |F|) from milli-mass.

    * ``rho = grid / 1000`` (complex.glsl:23)
    * forward unnormalised DFT, sign ``-`` (fft_batched.glsl:25)
    * DC := 0 (fourier.glsl:16-18)
    * ``phi = rho_hat * (-2*pi*G / k^2)`` where float32 ``k2 > 1e-4``
    * ``grad_d = i * k_d * phi`` (``mult_ik``, greens.glsl:23, 53-60)
    * ``F_d = Re(unnormalised inverse DFT) / g^3`` (pack.glsl:20-27)
    * ``compat="reference"``: x slot := z force when ``aliases_x_with_z(g)``
      (the buffer aliasing of main.py:638-654); ``"physical"``: true x.

    ``fft="numpy"`` uses ``numpy.fft`` (FP64); ``fft="radix2"`` restates the
    reference's stage structure, optionally with its sub-1e-6 flush.

    ``solver="convolution"``: the reference runs the FFT solver whatever the
    key says (``step`` calls ``step_fft``, main.py:671-672), so with
    ``compat="reference"`` the result is the FFT path's; with
    ``compat="physical"`` it is ``solve_convolution``.
    """
    if solver == "convolution" and compat == "physical":
        return solve_convolution(grids, g, W, G)
    grids = np.asarray(grids)
    k = grids.shape[0]
    rho = grids.astype(np.float64) / 1000.0
    if fft == "numpy":
        rho_hat = np.fft.fftn(rho, axes=(1, 2, 3))
    else:
        rho_hat = fft3_radix2(rho, +1, flush)
    phi = rho_hat * greens_factor(g, W, G)[None]
    kv = wave_numbers(g, W)
    kx = kv[None, None, None, :]
    ky = kv[None, None, :, None]
    kz = kv[None, :, None, None]
    out = np.zeros((k, g, g, g, 4), dtype=np.float64)
    for d, kd in enumerate((kx, ky, kz)):
        grad = 1j * kd * phi
        if fft == "numpy":
            real = np.fft.ifftn(grad, axes=(1, 2, 3)).real
        else:
            real = fft3_radix2(grad, -1, flush).real / float(g**3)
        out[..., d] = real
    if compat == "reference" and aliases_x_with_z(g):
        out[..., 0] = out[..., 2]
    out[..., 3] = np.sqrt(out[..., 0] ** 2 + out[..., 1] ** 2 + out[..., 2] ** 2)
    return out


def convolution_kernel32(g, W):
    """Green's-gradient kernel of ``init_greens_kernel`` (main.py:250-278):
    ``K[iz, iy, ix, d] = r_d / (4 pi |r|^3)`` with the minimum-image offsets
    ``i if i <= g/2 else i - g`` in units of ``voxel = W / g``, zero at the
    origin, float32 values (the reference's ``"f4"`` array)."""
    voxel = W / g
    s = np.arange(g)
    s = np.where(s <= g // 2, s, s - g).astype(np.float64)
    dz, dy, dx = np.meshgrid(s, s, s, indexing="ij")
    rx, ry, rz = dx * voxel, dy * voxel, dz * voxel
    r2 = rx * rx + ry * ry + rz * rz
    r = np.sqrt(r2)
    with np.errstate(divide="ignore", invalid="ignore"):
        factor = np.where(r > 1e-10, 1.0 / (4.0 * np.pi * r2 * r), 0.0)
    K = np.zeros((g, g, g, 3), dtype=np.float32)
    K[..., 0] = rx * factor
    K[..., 1] = ry * factor
    K[..., 2] = rz * factor
    return K


def solve_convolution(grids, g, W, G, method="fft"):
    """``solver: convolution`` as a direct periodic real-space convolution of
    the density (mass per cell, ``grid / 1000``) with ``convolution_kernel32``:
    ``F_d(c) = 2 pi G voxel^3 sum_c' rho(c') K_d(c - c')`` -- the coupling of
    the spectral solve (grad phi with del^2 phi = 2 pi G rho), so a lone mass
    m gives ``G voxel^3 / 2 * m * r / |r|^3`` at every other cell.

    ``method="fft"`` evaluates the circular convolution with FP64 numpy.fft;
    ``method="direct"`` sums over all cell pairs (O(g^6), for small g).
    """
    grids = np.asarray(grids)
    k = grids.shape[0]
    rho = grids.astype(np.float64) / 1000.0
    K = convolution_kernel32(g, W).astype(np.float64)
    scale = 2.0 * np.pi * G * (W / g) ** 3
    out = np.zeros((k, g, g, g, 4), dtype=np.float64)
    if method == "fft":
        rho_hat = np.fft.fftn(rho, axes=(1, 2, 3))
        for d in range(3):
            k_hat = np.fft.fftn(K[..., d])
            out[..., d] = scale * np.fft.ifftn(rho_hat * k_hat[None], axes=(1, 2, 3)).real
    else:
        idx = np.arange(g)
        for i in range(k):
            for z, y, x in np.ndindex(g, g, g):
                dz = (z - idx)[:, None, None] % g
                dy = (y - idx)[None, :, None] % g
                dx = (x - idx)[None, None, :] % g
                out[i, z, y, x, :3] = scale * np.einsum("zyx,zyxd->d", rho[i], K[dz, dy, dx])
    out[..., 3] = np.sqrt(out[..., 0] ** 2 + out[..., 1] ** 2 + out[..., 2] ** 2)
    return out


def direct_dft3(x, direction):
    """Brute-force O(g^6) unnormalised 3D DFT over the last three axes."""
    x = np.asarray(x, dtype=np.complex128)
    g = x.shape[-1]
    j = np.arange(g)
    F = np.exp(-direction * 2j * np.pi * np.outer(j, j) / g)
    flat = x.reshape(x.shape[:-3] + (g**3,))
    zz, yy, xx = np.meshgrid(j, j, j, indexing="ij")
    zz, yy, xx = zz.ravel(), yy.ravel(), xx.ravel()
    M = F[zz[:, None], zz[None, :]] * F[yy[:, None], yy[None, :]] * F[xx[:, None], xx[None, :]]
    return (flat @ M.T).reshape(x.shape)


# ---------------------------------------------------------------------------
# A.5 Gather and update  (update.glsl:53-89)
# ---------------------------------------------------------------------------


def gather(pos, rot, off, force, P: Params, cells=None):
    """FP64 averaged world-frame force ``avgForce`` (n,3).

    ``worldForce = transpose(rot) * gridForce`` (update.glsl:77) with
    ``rot = mat3(col0,col1,col2)`` => ``world_m = sum_j rot[4m+j] * F_j``
    (= M F); summed over grids and divided by k (update.glsl:81).
    """
    c = centered32(pos, P)
    Ms = rotation_matrices(rot).astype(np.float64)
    acc = np.zeros((c.shape[0], 3), dtype=np.float64)
    for i in range(P.k):
        idx = cell_indices_one(c, rot[i], off[i], P) if cells is None else cells[i]
        F = force[i].reshape(-1, 4)[idx, :3].astype(np.float64)
        acc += F @ Ms[i].T
    return acc / P.k


def update32(pos, vel, avg_force, dt, damping, world_size, timescale=1.0):
    """Kick, damp, drift, wrap (update.glsl:83-89) in float32, no FMA.

    ``vel -= avgForce * deltaTime; vel *= damping; pos += vel * deltaTime;
    pos = mod(pos, worldSize)``; ``pos.w`` keeps the mass, ``vel.w = 0``.
    Returns new (pos, vel) as (n,4) float32 arrays.
    """
    pos = np.asarray(pos, dtype=np.float32)
    vel = np.asarray(vel, dtype=np.float32)
    a = np.asarray(avg_force, dtype=np.float32)[:, :3]
    dtp = F32(float(dt) * float(timescale))
    damp = F32(damping)
    W = F32(world_size)
    v = vel[:, :3]
    v = (v - (a * dtp).astype(np.float32)).astype(np.float32)
    v = (v * damp).astype(np.float32)
    x = (pos[:, :3] + (v * dtp).astype(np.float32)).astype(np.float32)
    x = glsl_mod32(x, W)
    out_p = np.empty_like(pos)
    out_v = np.zeros_like(vel)
    out_p[:, :3] = x
    out_p[:, 3] = pos[:, 3]
    out_v[:, :3] = v
    return out_p, out_v


# ---------------------------------------------------------------------------
# Whole step
# ---------------------------------------------------------------------------


def forces(pos, rot, off, n_grids, grid_size, world_size, G, compat="reference", fft="numpy", flush=False,
           solver="fft"):
    """FP64 ``avgForce`` (n,3) of the whole MOG pipeline, processed grid by
    grid (each grid's solve depends only on its own density), O(n) memory.

    Returns ``(avg_force, grids)`` with the int64 density grids.
    """
    P = Params.make(n_grids, grid_size, world_size)
    g = P.g
    c = centered32(pos, P)
    milli = milli_mass(pos)
    uniform = milli.size > 0 and bool(np.all(milli == milli[0]))
    Ms = rotation_matrices(rot).astype(np.float64)
    acc = np.zeros((c.shape[0], 3), dtype=np.float64)
    grids = np.zeros((P.k, g, g, g), dtype=np.int64)
    for i in range(P.k):
        idx = cell_indices_one(c, rot[i], off[i], P)
        if uniform:
            gi = np.bincount(idx, minlength=g**3).astype(np.int64) * (milli[0] if milli.size else 0)
        else:
            gi = _bincount_i64(idx, milli, g**3)
        grids[i] = gi.reshape(g, g, g)
        F = solve(grids[i : i + 1], g, world_size, G, compat=compat, fft=fft, flush=flush, solver=solver)[0]
        acc += F.reshape(-1, 4)[idx, :3] @ Ms[i].T
    return acc / P.k, grids


def step(pos, vel, rot, off, cfg, compat="reference", fft="numpy", flush=False):
    """One reference step: FP64 force, float32 state update.

    ``cfg`` holds n_grids, grid_size, world_size, G, dt, damping and
    optionally timescale.  Returns ``(pos, vel, avg_force)``.
    """
    avg, _ = forces(
        pos, rot, off, cfg["n_grids"], cfg["grid_size"], cfg["world_size"], cfg["G"], compat, fft, flush
    )
    p, v = update32(
        pos, vel, avg, cfg["dt"], cfg.get("damping", 1.0), cfg["world_size"], cfg.get("timescale", 1.0)
    )
    return p, v, avg


# ---------------------------------------------------------------------------
# FP64 diagnostics (no reference counterpart: the reference prints only its
# configuration, main.py:147-155)
# ---------------------------------------------------------------------------


def potential_energy_grids(grids, g, W, G):
    """FP64 potential energy of each grid's density, (k,) array.

    With ``rho = grid / 1000`` (mass per cell, complex.glsl:23) and the
    potential ``phi`` of the step's spectral Green's function
    (``phi_hat = G_hat rho_hat``, ``G_hat = -2 pi G / k^2`` where float32
    ``k2 > 1e-4``, DC zeroed; greens.glsl:40-47), the energy
    ``U = 1/2 sum_c rho(c) phi(c)`` equals, by Parseval for the
    unnormalised DFT, ``1/(2 g^3) sum_k G_hat(k) |rho_hat(k)|^2``.  The force
    grids of ``solve`` are ``grad phi`` of this same ``phi``.
    """
    rho = np.asarray(grids).astype(np.float64) / 1000.0
    rho_hat = np.fft.fftn(rho, axes=(1, 2, 3))
    gf = greens_factor(g, W, G)
    return 0.5 / float(g**3) * np.einsum("kzyx,zyx->k", np.abs(rho_hat) ** 2, gf)


def diagnostics(pos, vel, rot, off, n_grids, grid_size, world_size, G):
    """FP64 momentum, kinetic and potential energy of a particle state.

    ``momentum = sum m v`` (3,), ``kinetic = 1/2 sum m |v|^2``,
    ``potential`` = mean over the k grids of ``potential_energy_grids`` of
    their NGP density (the energy whose gradient the k-averaged force
    approximates: the step accelerates by ``-avgForce``, update.glsl:83),
    ``mass = sum m``.  Masses are the float32 ``pos.w``; the potential uses
    the deposited milli-mass, like the force.
    """
    p = np.asarray(pos, dtype=np.float32)
    v = np.asarray(vel, dtype=np.float32)[:, :3].astype(np.float64)
    m = p[:, 3].astype(np.float64)
    P = Params.make(n_grids, grid_size, world_size)
    grids = deposit(p, rot, off, P)
    U = potential_energy_grids(grids, grid_size, world_size, G)
    return {
        "momentum": (m[:, None] * v).sum(axis=0),
        "kinetic": 0.5 * float(np.sum(m * np.sum(v * v, axis=1))),
        "potential": float(np.mean(U)),
        "mass": float(m.sum()),
    }


def rel_rms(a, ref):
    """sqrt(sum |a - ref|^2 / sum |ref|^2) over particles."""
    a = np.asarray(a, dtype=np.float64)[:, :3]
    ref = np.asarray(ref, dtype=np.float64)[:, :3]
    den = float(np.sum(ref * ref))
    num = float(np.sum((a - ref) ** 2))
    return math.sqrt(num / den) if den > 0 else math.sqrt(num)
