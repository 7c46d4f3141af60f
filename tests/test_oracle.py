"""Pins of the CPU oracle (no GPU).

Each check is independent of the formula it pins: a library routine, a
closed form, an invariant, a literal emulation of the reference's GLSL or
buffer-naming semantics, or brute force on small inputs.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import mog_oracle as O

W, G = 100.0, 100.0


def packed_identity(k=1):
    rot = np.zeros((k, 12), np.float32)
    rot[:, 0] = rot[:, 5] = rot[:, 10] = 1.0
    return rot, np.zeros((k, 4), np.float32)


def packed_from_matrix(M, offset=(0.0, 0.0, 0.0)):
    rot = np.zeros((1, 12), np.float32)
    for j in range(3):
        rot[0, 4 * j : 4 * j + 3] = M[j]
    off = np.zeros((1, 4), np.float32)
    off[0, :3] = offset
    return rot, off


# --- transforms (main.py:17-62, 157-172) -----------------------------------


def test_quaternion_matrix_matches_scipy():
    rng = np.random.default_rng(0)
    q = rng.normal(size=(200, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    for qi in q.astype(np.float32):
        M = O.quat_to_matrix(qi)  # (w, x, y, z)
        R = Rotation.from_quat([qi[1], qi[2], qi[3], qi[0]]).as_matrix()  # scalar-last
        assert np.abs(M - R).max() < 2e-6


def test_quaternion_known_rotation():
    th = np.pi / 2
    M = O.quat_to_matrix([np.cos(th / 2), 0, 0, np.sin(th / 2)])
    assert np.allclose(M, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)


@pytest.mark.parametrize("k,g,Wv,seed", [(240, 16, 100.0, 0), (16, 8, 2000.0, 3), (7, 32, 10.0, 11)])
def test_transforms_invariants_and_draw_order(k, g, Wv, seed):
    rot, off = O.se3_transforms(k, g, Wv, seed)
    M = O.rotation_matrices(rot).astype(np.float64)
    assert np.abs(np.einsum("kij,klj->kil", M, M) - np.eye(3)).max() < 1e-6
    assert np.allclose(np.linalg.det(M), 1.0, atol=1e-6)
    assert np.all(rot[:, [3, 7, 11]] == 0) and np.all(off[:, 3] == 0)
    h = Wv / g
    assert np.all(off[:, :3] >= 0) and np.all(off[:, :3] < h)
    rng = np.random.default_rng(seed)
    rng.uniform(0, 1, (k, 3))
    assert np.array_equal(off[:, :3], rng.uniform(0, h, (k, 3)).astype(np.float32))


def test_transforms_haar_statistics():
    rot, _ = O.se3_transforms(20000, 16, 100.0, 5)
    M = O.rotation_matrices(rot).astype(np.float64)
    assert np.abs(M.mean(0)).max() < 0.02
    assert np.abs((M**2).mean(0) - 1.0 / 3.0).max() < 0.02


def test_glsl_column_major_semantics():
    """``mat3(col0, col1, col2) * v`` with col_c = rotations[3i+c].xyz
    (mass.glsl:20-25): result = col0*v.x + col1*v.y + col2*v.z."""
    rot, off = O.se3_transforms(5, 16, W, 1)
    rng = np.random.default_rng(2)
    v = rng.uniform(-50, 50, (3,))
    for i in range(5):
        cols = [rot[i, 4 * c : 4 * c + 3].astype(np.float64) for c in range(3)]
        glsl = cols[0] * v[0] + cols[1] * v[1] + cols[2] * v[2]
        M = O.rotation_matrices(rot[i : i + 1])[0].astype(np.float64)
        assert np.allclose(glsl, M.T @ v, atol=1e-9)


# --- cell index (mass.glsl:42-50) ------------------------------------------


@pytest.mark.parametrize("g", [4, 16, 32])
def test_cell_identity_known_answers(g):
    P = O.Params.make(1, g, W)
    h = W / g
    idx = np.arange(g)
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    pos = np.zeros((g**3, 4), np.float32)
    pos[:, 0] = (xx.ravel() + 0.5) * h
    pos[:, 1] = (yy.ravel() + 0.5) * h
    pos[:, 2] = (zz.ravel() + 0.5) * h
    rot, off = packed_identity()
    cells = O.cell_indices(pos, rot, off, P)[0]
    assert np.array_equal(cells, (zz.ravel() * g + yy.ravel()) * g + xx.ravel())


def test_cell_offset_and_wrap():
    g = 16
    P = O.Params.make(1, g, W)
    rot, off = packed_identity()
    off[0, :3] = [1.0, 2.0, 3.0]
    pos = np.array([[0.5, 0.5, 0.5, 1.0], [99.9, 50.0, 10.0, 1.0], [150.0, -10.0, 205.0, 1.0]], np.float32)
    cells = O.cell_indices(pos, rot, off, P)[0]
    h = W / g
    exp = []
    for p in pos[:, :3].astype(np.float64):
        l = np.mod(p - [1.0, 2.0, 3.0], W)
        c = np.floor(l / h).astype(int) % g
        exp.append((c[2] * g + c[1]) * g + c[0])
    assert list(cells) == exp


def test_cell_rotation_known_answer():
    """90 degrees about z: local = M^T c maps (a,0,0) to (0,-a,0)."""
    M = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    rot, off = packed_from_matrix(M)
    P = O.Params.make(1, 16, W)
    pos = np.array([[60.0, 50.0, 50.0, 1.0]], np.float32)
    cell = O.cell_indices(pos, rot, off, P)[0, 0]
    x, y, z = cell % 16, (cell // 16) % 16, cell // 256
    assert (x, y, z) == (8, int(40.0 // 6.25), 8)


def test_cell_float32_matches_fp64_away_from_faces():
    g, k = 16, 12
    P = O.Params.make(k, g, W)
    rot, off = O.se3_transforms(k, g, W, 4)
    rng = np.random.default_rng(8)
    pos = np.zeros((20000, 4), np.float32)
    pos[:, :3] = rng.uniform(-20, 120, (20000, 3))
    cells = O.cell_indices(pos, rot, off, P)
    Ms = O.rotation_matrices(rot).astype(np.float64)
    c = pos[:, :3].astype(np.float64) - W / 2
    h = W / g
    for i in range(k):
        l = c @ Ms[i] + W / 2 - off[i, :3].astype(np.float64)  # (M^T c)^T = c^T M
        l = np.mod(l, W)
        q = l / h
        ref = np.floor(q).astype(int) % g
        flat = (ref[:, 2] * g + ref[:, 1]) * g + ref[:, 0]
        frac = np.abs(q - np.round(q))
        safe = np.all(frac > 1e-4, axis=1) & np.all(np.abs(l - W) > 1e-3, axis=1)
        assert safe.mean() > 0.99
        assert np.array_equal(cells[i][safe], flat[safe])


# --- deposit (mass.glsl:52-54) ---------------------------------------------


def test_milli_mass_truncation():
    pos = np.zeros((5, 4), np.float32)
    pos[:, 3] = [1.0, 1.0019999, 0.0015, -2.0, 2.5]
    m = O.milli_mass(pos)
    assert list(m) == [1000, int(np.float32(np.float32(1.0019999) * np.float32(1000))), 1, 0, 2500]


def test_deposit_conservation_permutation_bruteforce():
    k, g = 6, 8
    P = O.Params.make(k, g, W)
    rot, off = O.se3_transforms(k, g, W, 2)
    rng = np.random.default_rng(3)
    n = 3000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, W, (n, 3))
    pos[:, 3] = rng.uniform(0, 4, n)
    grids = O.deposit(pos, rot, off, P)
    milli = O.milli_mass(pos)
    assert np.all(grids.reshape(k, -1).sum(1) == milli.sum())
    perm = rng.permutation(n)
    assert np.array_equal(O.deposit(pos[perm], rot, off, P), grids)
    cells = O.cell_indices(pos, rot, off, P)
    brute = np.zeros((k, g**3), np.int64)
    for i in range(k):
        for p in range(n):
            brute[i, cells[i, p]] += milli[p]
    assert np.array_equal(brute.reshape(grids.shape), grids)


# --- solve (complex/fft/fourier/greens/pack.glsl) ---------------------------


def test_textbook_and_radix2_transforms_match_direct_dft():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 4, 4, 4)) + 1j * rng.normal(size=(2, 4, 4, 4))
    d = O.direct_dft3(x, +1)
    assert np.abs(np.fft.fftn(x, axes=(1, 2, 3)) - d).max() < 1e-12
    assert np.abs(O.fft3_radix2(x, +1) - d).max() < 1e-12
    di = O.direct_dft3(x, -1)
    assert np.abs(O.fft3_radix2(x, -1) - di).max() < 1e-12
    y = rng.normal(size=(1, 16, 16, 16))
    assert np.abs(O.fft3_radix2(y, +1) - np.fft.fftn(y, axes=(1, 2, 3))).max() < 1e-9


@pytest.mark.parametrize("g,axis", [(16, 0), (16, 1), (8, 2), (32, 0)])
def test_single_mode_closed_form(g, axis):
    """rho = A cos(k x_d) gives grad phi = 2 pi G A sin(k x_d) / k."""
    nmode, A = 3, 2.5
    j = np.arange(g)
    shape = [1, 1, 1]
    shape[2 - axis] = g
    wave = np.cos(2 * np.pi * nmode * j / g).reshape(shape)
    milli = np.broadcast_to(1000.0 * A * wave, (g, g, g))[None].astype(np.float64)
    F = O.solve(milli, g, W, G, compat="physical")[0]
    kk = 2 * np.pi * nmode / W
    expect = (2 * np.pi * G * A * np.sin(2 * np.pi * nmode * j / g) / kk).reshape(shape)
    expect = np.broadcast_to(expect, (g, g, g))
    assert np.abs(F[..., axis] - expect).max() < 1e-9 * np.abs(expect).max()
    for other in {0, 1, 2} - {axis}:
        assert np.abs(F[..., other]).max() < 1e-9 * np.abs(expect).max()
    Fr = O.solve(milli, g, W, G, compat="reference")[0]
    if O.aliases_x_with_z(g):
        assert np.array_equal(Fr[..., 0], Fr[..., 2])
    else:
        assert np.array_equal(Fr, F)


def _reference_buffer_names(g):
    """Literal restatement of main.py's texture naming: which textures pack
    reads as (x, y, z) after step_fft (main.py:564-590, 617-654)."""
    stages = int(math.log2(g))

    def run(src, dst):
        cur_src, cur_dst = src, dst
        for _axis in range(3):
            for _stage in range(stages + 1):
                cur_src, cur_dst = cur_dst, cur_src
        return cur_src

    spectrum = run("a", "b")
    contents = {"a": None, "b": None, "gx": "gx_spec", "gy": "gy_spec", "gz": "gz_spec"}
    del spectrum
    out = []
    for grad, scratch, name in (("gx", "b", "X"), ("gy", "a", "Y"), ("gz", "b", "Z")):
        res = run(grad, scratch)
        contents[res] = name
        out.append(res)
    return tuple(contents[t] for t in out)


@pytest.mark.parametrize("g", [2, 4, 8, 16, 32, 64])
def test_reference_aliasing_matches_buffer_logic(g):
    read = _reference_buffer_names(g)
    if O.aliases_x_with_z(g):
        assert read == ("Z", "Y", "Z")
    else:
        assert read == ("X", "Y", "Z")


def test_odd_kernel_invariants_and_pair_symmetry():
    g = 16
    rng = np.random.default_rng(6)
    milli = rng.integers(0, 5000, (1, g, g, g)).astype(np.int64)
    F = O.solve(milli, g, W, G, compat="physical")[0]
    rho = milli[0] / 1000.0
    for d in range(3):
        s = np.sum(rho * F[..., d])
        assert abs(s) < 1e-10 * np.sum(np.abs(rho) * np.abs(F[..., d]))
    two = np.zeros((1, g, g, g), np.int64)
    two[0, 3, 4, 5] = 1000
    two[0, 9, 7, 2] = 1000
    F2 = O.solve(two, g, W, G, compat="physical")[0]
    assert np.allclose(F2[3, 4, 5, :3], -F2[9, 7, 2, :3], rtol=1e-10, atol=1e-10)
    # attraction: grad phi at particle 1 points to particle 2 (update subtracts it)
    d12 = np.array([2 - 5, 7 - 4, 9 - 3], float)
    assert np.dot(F2[3, 4, 5, :3], d12) < 0


def test_greens_mask_large_world():
    """At W=2000 the float32 k2 > 1e-4 test drops every mode with
    nx^2+ny^2+nz^2 <= 10 (greens.glsl:45)."""
    g = 16
    mask = O.greens_mask32(g, 2000.0)
    s = np.array([i if i <= g // 2 else i - g for i in range(g)])
    n2 = s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2
    assert np.array_equal(mask, n2 > 10)
    assert np.array_equal(O.greens_mask32(g, 100.0), n2 > 0)


def test_flush_variant_effect_is_below_tolerance():
    g = 16
    rng = np.random.default_rng(12)
    milli = rng.integers(0, 20000, (2, g, g, g)).astype(np.int64)
    a = O.solve(milli, g, W, G, fft="radix2", flush=False)
    b = O.solve(milli, g, W, G, fft="radix2", flush=True)
    c = O.solve(milli, g, W, G, fft="numpy")
    assert O.rel_rms(a[..., :3].reshape(-1, 3), c[..., :3].reshape(-1, 3)) < 1e-13
    assert O.rel_rms(b[..., :3].reshape(-1, 3), c[..., :3].reshape(-1, 3)) < 1e-8


# --- gather / update (update.glsl) ------------------------------------------


def test_gather_identity_and_rotated_uniform_field():
    g = 8
    P1 = O.Params.make(1, g, W)
    rot, off = packed_identity()
    force = np.zeros((1, g, g, g, 4))
    force[..., :3] = [1.0, 2.0, 3.0]
    rng = np.random.default_rng(1)
    pos = np.zeros((100, 4), np.float32)
    pos[:, :3] = rng.uniform(0, W, (100, 3))
    assert np.allclose(O.gather(pos, rot, off, force, P1), [1.0, 2.0, 3.0])
    M = O.quat_to_matrix([np.cos(0.3), np.sin(0.3) * 0.6, 0.0, np.sin(0.3) * 0.8])
    rot2, off2 = packed_from_matrix(M, (0.5, 1.0, 2.0))
    got = O.gather(pos, rot2, off2, force, P1)
    assert np.allclose(got, M.astype(np.float64) @ np.array([1.0, 2.0, 3.0]), atol=1e-6)


def test_update_zero_force_closed_form():
    rng = np.random.default_rng(2)
    pos = np.zeros((1000, 4), np.float32)
    pos[:, :3] = rng.uniform(0, W, (1000, 3))
    pos[:, 3] = 1.5
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 1e4, (1000, 3))
    p, v = O.update32(pos, vel, np.zeros((1000, 3)), 1e-3, 0.9, W)
    v64 = vel[:, :3].astype(np.float64) * 0.9
    x64 = np.mod(pos[:, :3].astype(np.float64) + v64 * 1e-3, W)
    assert np.allclose(v[:, :3], v64, rtol=1e-6)
    d = np.abs(p[:, :3] - x64)
    assert np.all(np.minimum(d, W - d) < 1e-4)
    assert np.all(p[:, 3] == 1.5) and np.all(v[:, 3] == 0)


def test_forces_pipeline_momentum_single_grid_pm():
    """k=1, identity transform reduces MOG to textbook single-grid PM:
    the mass-weighted sum of forces vanishes (physical mode)."""
    g = 16
    rot, off = packed_identity()
    rng = np.random.default_rng(3)
    pos = np.zeros((5000, 4), np.float32)
    pos[:, :3] = rng.uniform(0, W, (5000, 3))
    pos[:, 3] = 1.0
    avg, _ = O.forces(pos, rot, off, 1, g, W, G, compat="physical")
    assert np.abs(avg.sum(0)).max() < 1e-9 * np.abs(avg).sum()


# --- solver: convolution ----------------------------------------------------


@pytest.mark.parametrize("g", [4, 8])
def test_convolution_matches_brute_force_pair_sum(g):
    rng = np.random.default_rng(g)
    grids = rng.integers(0, 5000, (2, g, g, g)).astype(np.int64)
    a = O.solve_convolution(grids, g, W, G, method="fft")
    b = O.solve_convolution(grids, g, W, G, method="direct")
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def test_convolution_lone_mass_closed_form():
    """A lone mass m at cell c0 gives G voxel^3 / 2 * m r / |r|^3 with r the
    minimum-image offset (Nyquist offset positive) -- Newton with the MOG
    coupling, no periodic images, no background."""
    g = 16
    voxel = W / g
    grids = np.zeros((1, g, g, g), np.int64)
    c0 = (3, 9, 14)  # (z, y, x)
    grids[0][c0] = 7000  # m = 7
    F = O.solve_convolution(grids, g, W, G)[0]
    for z, y, x in [(3, 9, 15), (0, 0, 0), (11, 1, 6), (7, 12, 2)]:
        off = [(c - c_0) % g for c, c_0 in zip((x, y, z), (c0[2], c0[1], c0[0]))]
        r = np.array([o if o <= g // 2 else o - g for o in off], dtype=np.float64) * voxel
        want = G * voxel**3 / 2 * 7.0 * r / np.linalg.norm(r) ** 3
        np.testing.assert_allclose(F[z, y, x, :3], want, rtol=1e-6, atol=1e-9 * np.max(np.abs(want)))
    assert np.max(np.abs(F[c0][:3])) <= 1e-9 * np.max(np.abs(F[..., :3]))


def test_convolution_zero_net_momentum_for_compact_density():
    """sum_c rho(c) F(c) = 0 when the density spans less than half the box
    (every pair offset then avoids the unpaired Nyquist plane, so the kernel
    is exactly odd on the pairs that occur)."""
    g = 16
    rng = np.random.default_rng(5)
    grids = np.zeros((1, g, g, g), np.int64)
    grids[0, 2:9, 4:10, 5:12] = rng.integers(0, 3000, (7, 6, 7))
    F = O.solve_convolution(grids, g, W, G)[0]
    rho = grids[0].astype(np.float64)[..., None] / 1000.0
    p = np.sum(rho * F[..., :3], axis=(0, 1, 2))
    assert np.max(np.abs(p)) <= 1e-10 * np.sum(np.abs(rho * F[..., :3]))


def test_convolution_key_semantics_and_agreement_with_fft():
    """compat reference: the key runs the FFT solver (as the reference does);
    compat physical: the real-space convolution, which agrees in direction
    and within 15% in magnitude with the spectral force of a lone mass at
    4-8 cells off the axes."""
    g = 32
    grids = np.zeros((1, g, g, g), np.int64)
    grids[0, 16, 16, 16] = 50_000
    ref = O.solve(grids, g, W, G, compat="reference", solver="convolution")
    np.testing.assert_array_equal(ref, O.solve(grids, g, W, G, compat="reference"))
    conv = O.solve(grids, g, W, G, compat="physical", solver="convolution")[0]
    np.testing.assert_array_equal(conv, O.solve_convolution(grids, g, W, G)[0])
    fft = O.solve(grids, g, W, G, compat="physical")[0]
    for d in (4, 6, 8):
        a, b = conv[16 + d // 2, 16 + d, 16, :3], fft[16 + d // 2, 16 + d, 16, :3]
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999
        assert 0.85 < np.linalg.norm(a) / np.linalg.norm(b) < 1.15


# ---------------------------------------------------------------------------
# Kick sign and the FP64 diagnostics (update.glsl:83-86; the step
# accelerates by -avgForce, avgForce = grad phi)
# ---------------------------------------------------------------------------


def two_clusters(n=8192, sigma=6.0, sep=40.0, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.normal([W / 2 - sep / 2, W / 2, W / 2], sigma, (n // 2, 3))
    b = rng.normal([W / 2 + sep / 2, W / 2, W / 2], sigma, (n // 2, 3))
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.mod(np.concatenate([a, b]), W)
    pos[:, 3] = 1.0
    return pos, np.zeros_like(pos)


def _infall(update=None, steps=20, dt=1e-4):
    """Two cold clusters under the oracle step (physical compat, no
    damping): (separation before, after, diagnostics before, after)."""
    k, g, Gc = 16, 16, 100.0
    pos, vel = two_clusters()
    rot, off = O.se3_transforms(k, g, W, 0)
    cfg = dict(n_grids=k, grid_size=g, world_size=W, G=Gc, dt=dt, damping=1.0)
    half = pos.shape[0] // 2
    sep0 = pos[half:, 0].mean() - pos[:half, 0].mean()
    d0 = O.diagnostics(pos, vel, rot, off, k, g, W, Gc)
    p, v = pos, vel
    for _ in range(steps):
        avg, _ = O.forces(p, rot, off, k, g, W, Gc, compat="physical")
        p, v = (update or O.update32)(p, v, avg, dt, 1.0, W)
    d1 = O.diagnostics(p, v, rot, off, k, g, W, Gc)
    return sep0, p[half:, 0].mean() - p[:half, 0].mean(), d0, d1


def test_kick_sign_infall_and_energy_exchange():
    """Two cold clusters fall towards each other, the kinetic energy they
    gain is paid by the potential energy (|dE| < 20 % of dKE; NGP grids do
    not conserve energy exactly), and the momentum stays zero.  A '+=' kick
    (repulsion) fails every one of these; the planted mistake below shows
    it."""
    sep0, sep1, d0, d1 = _infall()
    assert sep1 < sep0 - 0.05
    dke = d1["kinetic"] - d0["kinetic"]
    de = (d1["kinetic"] + d1["potential"]) - (d0["kinetic"] + d0["potential"])
    assert dke > 0 and abs(de) < 0.2 * dke, (de, dke)
    vrms = math.sqrt(2 * d1["kinetic"] / d1["mass"])
    assert np.linalg.norm(d1["momentum"]) < 1e-6 * d1["mass"] * vrms


def test_flipped_kick_is_caught():
    """A planted '+=' kick in a copy of the oracle's integrator: the clusters
    separate and the energy is not conserved."""

    def update_plus(pos, vel, avg, dt, damping, world_size, timescale=1.0):
        return O.update32(pos, vel, -np.asarray(avg), dt, damping, world_size, timescale)

    sep0, sep1, d0, d1 = _infall(update=update_plus)
    dke = d1["kinetic"] - d0["kinetic"]
    de = (d1["kinetic"] + d1["potential"]) - (d0["kinetic"] + d0["potential"])
    assert sep1 > sep0
    assert abs(de) > 0.5 * dke


def test_potential_energy_parseval_and_single_mode():
    """potential_energy_grids against (i) 1/2 sum rho phi with phi from a
    real-space inverse transform (an independent path through the same
    Green's function) and (ii) the closed form of one Fourier mode,
    rho = A cos(2 pi n x / W): U = G_hat(k) A^2 g^3 / 4."""
    g, Gc = 16, 100.0
    rng = np.random.default_rng(5)
    grids = rng.integers(0, 5000, (3, g, g, g)).astype(np.int64)
    U = O.potential_energy_grids(grids, g, W, Gc)
    rho = grids.astype(np.float64) / 1000.0
    phi = np.fft.ifftn(np.fft.fftn(rho, axes=(1, 2, 3)) * O.greens_factor(g, W, Gc)[None], axes=(1, 2, 3)).real
    direct = 0.5 * np.sum(rho * phi, axis=(1, 2, 3))
    assert np.allclose(U, direct, rtol=1e-12)
    assert np.all(U < 0)
    A, nx = 2.5, 3
    x = np.arange(g)
    rho1 = np.broadcast_to(A * np.cos(2 * np.pi * nx * x / g), (g, g, g))
    U1 = O.potential_energy_grids(1000.0 * rho1[None], g, W, Gc)[0]
    kx = 2 * np.pi * nx / W
    assert np.isclose(U1, -2 * np.pi * Gc / kx**2 * A**2 * g**3 / 4, rtol=1e-12)
