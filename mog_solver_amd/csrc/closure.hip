// Learned closure of the MOG step: features of a cheap k-grid ensemble and a
// per-particle MLP that predicts the force of a larger ensemble.
//
// Features (per particle, from a context with few grids, e.g. k = 24):
//   m   = mean over grids of the world-frame force M_i F_i(cell), NGP    (3)
//   s   = across-grid standard deviation of those forces, per axis      (3)
//   c   = mean over grids of the trilinear (cloud-in-cell) interpolation
//         of the deconvolved world-frame force field: each grid's force
//         spectrum divided by the NGP window sinc(k_x h/2) sinc(k_y h/2)
//         sinc(k_z h/2), sampled at the particle's continuous cell
//         coordinate.  The average of NGP forces over many random offsets
//         is the field smoothed by one cell; interpolating the deconvolved
//         field estimates that average from few grids, so c is a much
//         better base than m (both are inputs)                          (3)
//   rho = mean over grids of the mass of the particle's cell            (1)
// Network input x = (m/S, s/S, c/S, log(1 + rho/rho0)), S = |c| + |s| + tiny
// (a per-particle force scale, so the network is scale-free), output y (3);
// predicted force F = c + S y.  feat holds 12 floats per particle: x[0..9],
// S, 0.  Layers 10 -> 64 -> 64 -> 64 -> 3 with SiLU activations;
// parameters fp32, row-major W[in][out], each followed by its bias:
// W1 b1 W2 b2 W3 b3 W4 b4 (9219 floats).
//
// GEMMs run on the matrix cores over tiles of 64 particles held in LDS:
// BF16 operands (v_mfma_f32_16x16x32_bf16) or FP32 operands
// (v_mfma_f32_16x16x4_f32), FP32 accumulation either way.  The training
// kernel is fused: forward (pre-activations kept in LDS), loss
// sum_p w_p |y_p - t_p|^2 with t = (F_target - c) / S and w = S^2 / mean S^2,
// backward, and gradient accumulation in LDS, flushed once per workgroup.
#include "mog_common.h"

namespace mog {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kIn = 10, kH = 64, kOut = 3;
constexpr int kXS = 16;  // LDS row stride of a tile's inputs
constexpr int oW1 = 0, ob1 = oW1 + kIn * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH,
              oW3 = ob2 + kH, ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + kH * kOut,
              kParams = ob4 + kOut;
constexpr int TILE = 64;  // particles per tile

// sigmoid with the precise exp and IEEE division (FP32 operands), or with
// the hardware exp2 and reciprocal, about 1 ulp each (BF16 operands: the
// activation is rounded to bf16 next).  The builtins are written out because
// this library is compiled with correctly rounded division, which would
// expand a quotient into the full IEEE sequence.
template <bool FAST>
__device__ __forceinline__ float sigm(float z) {
  if (FAST) return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * z));
  return 1.0f / (1.0f + expf(-z));
}
template <bool FAST>
__device__ __forceinline__ float silu(float z) {
  return z * sigm<FAST>(z);
}
template <bool FAST>
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm<FAST>(z);
  return s * (1.0f + z * (1.0f - s));
}

// A matrix view: element (r, c) at p[r * rs + c * cs]; reads outside
// (nr, nc) are zero; act applies SiLU on read.
struct Mat {
  const float* p;
  int rs, cs, nr, nc;
  bool act;
  template <bool FAST>
  __device__ float at(int r, int c) const {
    if (r >= nr || c >= nc) return 0.0f;
    const float v = p[r * rs + c * cs];
    return act ? silu<FAST>(v) : v;
  }
};

// C[M x N] (+)= A[M x K] . B[K x N] by the workgroup's waves; M, N are
// multiples of 16 (padding reads zero), output written only inside
// (cr, cc).  BF16: K handled in chunks of 32; FP32: chunks of 4.
template <bool BF16>
__device__ void wg_gemm(const Mat A, const Mat B, float* C, int ldc, int M, int N, int K, int cr,
                        int cc, bool accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int tn = N / 16, nt = (M / 16) * tn;
  for (int t = wave; t < nt; t += nw) {
    const int r0 = (t / tn) * 16, c0 = (t % tn) * 16;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const int ar = r0 + (lane & 15), bc = c0 + (lane & 15);
    if (BF16) {
      for (int k0 = 0; k0 < K; k0 += 32) {
        v8bf a, b;
        const int kb = k0 + 8 * (lane >> 4);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          a[j] = (__bf16)A.template at<true>(ar, kb + j);
          b[j] = (__bf16)B.template at<true>(kb + j, bc);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
      }
    } else {
      for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + (lane >> 4);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A.template at<false>(ar, k),
                                                  B.template at<false>(k, bc), acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 4 * (lane >> 4) + i;
      if (r < cr && bc < cc) {
        if (accumulate)
          C[r * ldc + bc] += acc[i];
        else
          C[r * ldc + bc] = acc[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Features of a k-grid ensemble (context grids and world-frame force grids)
// ---------------------------------------------------------------------------
// Per cell of the k context grids (one thread per cell): fm[c] = (F_c.xyz,
// mass of the cell x 1e-3, the value the features' rho sums), and the
// trilinear form of the deconvolved field D over the octant above the cell:
// with D_j at the corner j = (jx, jy, jz) = (j & 1, j >> 1 & 1, j >> 2)
// (cell c + jx x + jy y + jz z, periodic),
//   D(c + f) = a0 + a1 fx + a2 fy + a3 fz + a4 fx fy + a5 fx fz + a6 fy fz
//              + a7 fx fy fz,   f in [0, 1)^3,
// i.e. the sum over corners of D_j times the weights (fx or 1 - fx)(fy or
// 1 - fy)(fz or 1 - fz).  A cell's block holds a_i.xyz at float 3 i + m
// (24 floats) and 8 floats of padding: 128 bytes, so a particle reads its
// interpolation with six 16-byte loads from one address.
__global__ __launch_bounds__(256) void closure_grids_kernel(const float4* __restrict__ force,
                                                            const float4* __restrict__ force_dc,
                                                            const long long* __restrict__ grids,
                                                            long long total, int lg,
                                                            float4* __restrict__ fm,
                                                            float4* __restrict__ blk) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= total) return;
  const int gm = (1 << lg) - 1;
  const long long cells = 1ll << (3 * lg);
  const long long base = c & ~(cells - 1);
  const int cell = (int)(c & (cells - 1));
  const int x = cell & gm, y = (cell >> lg) & gm, z = cell >> (2 * lg);
  const float4 F = force[c];
  fm[c] = make_float4(F.x, F.y, F.z, (float)grids[c] * 1e-3f);
  float d[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int xx = (x + (j & 1)) & gm, yy = (y + ((j >> 1) & 1)) & gm, zz = (z + (j >> 2)) & gm;
    const float4 v = force_dc[base + ((((long long)zz << lg) | yy) << lg | xx)];
    d[j][0] = v.x;
    d[j][1] = v.y;
    d[j][2] = v.z;
  }
  float a[32];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    a[m] = d[0][m];
    a[3 + m] = d[1][m] - d[0][m];
    a[6 + m] = d[2][m] - d[0][m];
    a[9 + m] = d[4][m] - d[0][m];
    a[12 + m] = (d[3][m] - d[1][m]) - (d[2][m] - d[0][m]);
    a[15 + m] = (d[5][m] - d[1][m]) - (d[4][m] - d[0][m]);
    a[18 + m] = (d[6][m] - d[2][m]) - (d[4][m] - d[0][m]);
    a[21 + m] = ((d[7][m] - d[6][m]) - (d[5][m] - d[4][m])) - ((d[3][m] - d[2][m]) - (d[1][m] - d[0][m]));
  }
#pragma unroll
  for (int i = 24; i < 32; ++i) a[i] = 0.f;
  float4* out = blk + 8 * c;
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = make_float4(a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]);
}

// The trilinear form of one component at f (closure_grids_kernel), Horner in
// fz, then fy, then fx.
__device__ __forceinline__ float trilinear(float a0, float a1, float a2, float a3, float a4,
                                           float a5, float a6, float a7, float fx, float fy,
                                           float fz) {
  const float tx = fmaf(fmaf(a7, fz, a4), fy, fmaf(a5, fz, a1));
  const float t0 = fmaf(fmaf(a6, fz, a2), fy, fmaf(a3, fz, a0));
  return fmaf(tx, fx, t0);
}

// Per particle, over the k grids: the NGP cell of the fast path (the force
// and cell mass from fm), the continuous cell coordinate u = l/h - 1/2 (cell
// centres at integers: the fast row's w) and the interpolation block of the
// cell floor(u).  NGP pairs the fast path cannot decide are deferred as in
// the force gather: recorded in the lane's LDS slots, their exact cells
// evaluated after the grid loop.  SH: NGP and block offsets through the
// shifted cell fields (kShiftF4 and kShiftBlk); otherwise from indices.
template <bool SH>
__global__ __launch_bounds__(256) void closure_features_kernel(
    const float4* __restrict__ pos, long long n, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    const float4* __restrict__ fm, const float4* __restrict__ blk, Params P, float rho0,
    const int* __restrict__ order, float4* __restrict__ feat, float4* __restrict__ mean_out) {
  __shared__ unsigned recs[kGatherSlots * 256];
  const long long t = (long long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // visiting order (spatially sorted when given): neighbouring lanes, and
  // neighbouring workgroups of one XCD, read neighbouring cells
  const long long i = order ? (long long)order[t] : t;
  const float4 p = pos[i];
  const float c0 = p.x - P.half, c1 = p.y - P.half, c2 = p.z - P.half;
  const int pfast = particle_fast(c0, c1, c2, P.CB);
  FieldShift FS = field_shift(P.lg, 4), FB = field_shift(P.lg, 7);
  if constexpr (SH) {  // vector-register operands issue at the full rate
    FS.k0 = to_vgpr(FS.k0);
    FS.k1 = to_vgpr(FS.k1);
    FS.k2 = to_vgpr(FS.k2);
    FS.m0 = to_vgpr(FS.m0);
    FS.m01 = to_vgpr(FS.m01);
    FS.m012 = to_vgpr(FS.m012);
    FB.k0 = to_vgpr(FB.k0);
    FB.k1 = to_vgpr(FB.k1);
    FB.k2 = to_vgpr(FB.k2);
    FB.m0 = to_vgpr(FB.m0);
    FB.m01 = to_vgpr(FB.m01);
    FB.m012 = to_vgpr(FB.m012);
  }
  const int gm = P.g - 1;
  const bool wave_slow = __builtin_amdgcn_ballot_w64(!pfast) != 0;
  const int lane = threadIdx.x;
  // Sums are taken about a force of the first grid (shifted data), so the
  // variance does not cancel catastrophically when the spread is small.
  float bx = 0.f, by = 0.f, bz = 0.f;
  float sx = 0.f, sy = 0.f, sz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f, sm = 0.f;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  int nrec = 0;
  for (int g = 0; g < P.k; ++g) {
    const float* F = frot + 12 * g;
    const float w0 = fmaf(F[2], c2, fmaf(F[1], c1, fmaf(F[0], c0, F[3])));
    const float w1 = fmaf(F[6], c2, fmaf(F[5], c1, fmaf(F[4], c0, F[7])));
    const float w2 = fmaf(F[10], c2, fmaf(F[9], c1, fmaf(F[8], c0, F[11])));
    const float fl0 = floorf(w0), fl1 = floorf(w1), fl2 = floorf(w2);
    const float4* Fg = fm + (long long)g * P.cells;
    const float4* Bg = blk + (long long)g * P.cells * 8;
    unsigned o;
    float d;
    float4 a[6];
    if constexpr (SH) {
      d = cell_fast_o(c0, c1, c2, F, F[3], F[7], F[11], FS, 0u, &o);
      const float u0 = fl0 + FB.k0, u1 = fl1 + FB.k1, u2 = fl2 + FB.k2;
      unsigned b0, b1, b2;
      memcpy(&b0, &u0, 4);
      memcpy(&b1, &u1, 4);
      memcpy(&b2, &u2, 4);
      const unsigned ob = and_or(bit_select(bit_select(b0, b1, FB.m0), b2, FB.m01), FB.m012, 0u);
      const float4* B = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(Bg) + ob);
#pragma unroll
      for (int j = 0; j < 6; ++j) a[j] = B[j];
    } else {
      int cell;
      d = cell_fast_d(c0, c1, c2, F, F[3], F[7], F[11], P.lg, &cell);
      o = (unsigned)cell << 4;
      const long long cb = (((long long)((int)fl2 & gm) << P.lg | ((int)fl1 & gm)) << P.lg) |
                           ((int)fl0 & gm);
      const float4* B = Bg + 8 * cb;
#pragma unroll
      for (int j = 0; j < 6; ++j) a[j] = B[j];
    }
    float4 f = cell_load_b(Fg, o);
    if (g == 0) {  // any force of the particle's grids serves as the shift
      bx = f.x;
      by = f.y;
      bz = f.z;
    }
    if (wave_slow || __builtin_amdgcn_ballot_w64(!(d < P.lim)) != 0) {
      if (!(pfast & (d < P.lim))) {
        if (nrec < kGatherSlots) {  // adds nothing now: zero deviation, no mass
          recs[nrec * 256 + lane] = (unsigned)g;
          f = make_float4(bx, by, bz, 0.f);
        } else {
          const int ex =
              cell_exact_call(c0, c1, c2, rot + 12 * g, off + 4 * g, P.W, P.half, P.h, P.lg);
          f = cell_load(Fg, ex);
        }
        ++nrec;
      }
    }
    const float ex = f.x - bx, ey = f.y - by, ez = f.z - bz;
    sx += ex;
    sy += ey;
    sz += ez;
    qx = fmaf(ex, ex, qx);
    qy = fmaf(ey, ey, qy);
    qz = fmaf(ez, ez, qz);
    sm += f.w;
    const float fx = w0 - fl0, fy = w1 - fl1, fz = w2 - fl2;
    cx += trilinear(a[0].x, a[0].w, a[1].z, a[2].y, a[3].x, a[3].w, a[4].z, a[5].y, fx, fy, fz);
    cy += trilinear(a[0].y, a[1].x, a[1].w, a[2].z, a[3].y, a[4].x, a[4].w, a[5].z, fx, fy, fz);
    cz += trilinear(a[0].z, a[1].y, a[2].x, a[2].w, a[3].z, a[4].y, a[5].x, a[5].w, fx, fy, fz);
  }
  // the deferred NGP pairs
  const int mine = min(nrec, kGatherSlots);
  for (int r = 0; r < mine; ++r) {
    const int g = (int)recs[r * 256 + lane];
    const int ec = cell_exact_call(c0, c1, c2, rot + 12 * g, off + 4 * g, P.W, P.half, P.h, P.lg);
    const float4 f = cell_load(fm + (long long)g * P.cells, ec);
    const float ex = f.x - bx, ey = f.y - by, ez = f.z - bz;
    sx += ex;
    sy += ey;
    sz += ez;
    qx = fmaf(ex, ex, qx);
    qy = fmaf(ey, ey, qy);
    qz = fmaf(ez, ez, qz);
    sm += f.w;
  }
  const float ik = 1.0f / (float)P.k;
  const float ux = sx * ik, uy = sy * ik, uz = sz * ik;
  const float mx = bx + ux, my = by + uy, mz = bz + uz;
  const float vx = fmaxf(qx * ik - ux * ux, 0.f), vy = fmaxf(qy * ik - uy * uy, 0.f),
              vz = fmaxf(qz * ik - uz * uz, 0.f);
  const float dx = sqrtf(vx), dy = sqrtf(vy), dz = sqrtf(vz);
  cx *= ik;
  cy *= ik;
  cz *= ik;
  const float S = sqrtf(cx * cx + cy * cy + cz * cz) + sqrtf(vx + vy + vz) + 1e-30f;
  const float iS = 1.0f / S;
  feat[3 * i] = make_float4(mx * iS, my * iS, mz * iS, dx * iS);
  feat[3 * i + 1] = make_float4(dy * iS, dz * iS, cx * iS, cy * iS);
  feat[3 * i + 2] = make_float4(cz * iS, log1pf(sm * ik / rho0), S, 0.f);
  mean_out[i] = make_float4(cx, cy, cz, 0.f);
}

// ---------------------------------------------------------------------------
// MLP inference: F = c + S * MLP(x) (c: the features' base estimate, "mean")
// ---------------------------------------------------------------------------
template <bool BF16>
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* __restrict__ prm,
                                                          const float4* __restrict__ feat,
                                                          const float4* __restrict__ mean,
                                                          long long n, float4* __restrict__ out) {
  __shared__ float X[TILE * kXS];
  __shared__ float H0[TILE * kH];
  __shared__ float H1[TILE * kH];
  __shared__ float Y[TILE * 16];
  const long long ntile = (n + TILE - 1) / TILE;
  for (long long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long long base = tile * TILE;
    for (int i = threadIdx.x; i < TILE * kXS; i += blockDim.x) {
      const long long p = base + i / kXS;
      const int f = i % kXS;
      float v = 0.f;
      if (p < n && f < kIn) {
        const float4 a = feat[3 * p + f / 4];
        v = (&a.x)[f % 4];
      }
      X[i] = v;
    }
    __syncthreads();
    // layer 1: Z1 = X W1 + b1 (bias added after); H0 holds Z1
    wg_gemm<BF16>(Mat{X, kXS, 1, TILE, kIn, false}, Mat{prm + oW1, kH, 1, kIn, kH, false}, H0, kH,
                  TILE, kH, 32, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) H0[i] += prm[ob1 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{H0, kH, 1, TILE, kH, true}, Mat{prm + oW2, kH, 1, kH, kH, false}, H1, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) H1[i] += prm[ob2 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{H1, kH, 1, TILE, kH, true}, Mat{prm + oW3, kH, 1, kH, kH, false}, H0, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) H0[i] += prm[ob3 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{H0, kH, 1, TILE, kH, true}, Mat{prm + oW4, kOut, 1, kH, kOut, false}, Y, 16,
                  TILE, 16, kH, TILE, kOut, false);
    __syncthreads();
    for (int r = threadIdx.x; r < TILE; r += blockDim.x) {
      const long long p = base + r;
      if (p < n) {
        const float sc = feat[3 * p + 2].z;
        const float4 m = mean[p];
        out[p] = make_float4(fmaf(sc, Y[r * 16 + 0] + prm[ob4 + 0], m.x),
                             fmaf(sc, Y[r * 16 + 1] + prm[ob4 + 1], m.y),
                             fmaf(sc, Y[r * 16 + 2] + prm[ob4 + 2], m.z), 0.f);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Register-resident BF16 inference (the production forward pass)
// ---------------------------------------------------------------------------
// One wave evaluates 16 particles per iteration with every weight held as a
// v_mfma_f32_16x16x32_bf16 A-fragment in VGPRs (88 VGPRs) and the
// activations never leaving registers.  The product is taken transposed,
// H^T = W^T X^T: the weights are the A operand (rows = output features) and
// the particles the B/C columns, so lane l's accumulators hold features
// 16t + 4q + i (tile t, q = l/16, i = 0..3) of particle l % 16.  The next
// layer's B operand needs, in lane group q, contraction indices 8q + j of a
// 32-wide K chunk; the contraction order is free, so chunk c slot j of group
// q is defined to be feature kperm(c, q, j) = 16(2c + j/4) + 4q + j%4 --
// exactly the lane's own accumulator h[2c + j/4][j%4] -- and the weight
// fragments are loaded in the same order.  No LDS round trip, no shuffles.
// Layer 1's bias rides in the padded K slot 10 (input 1.0).
__device__ __forceinline__ int kperm(int c, int q, int j) {
  return 16 * (2 * c + (j >> 2)) + 4 * q + (j & 3);
}

// z * sigmoid(z) with the hardware exp2 and reciprocal (about 1 ulp each;
// the result is rounded to bf16 next).  Written with the builtins because
// this library is compiled with correctly rounded division, which would
// expand a quotient into the full IEEE sequence.
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}

__device__ __forceinline__ v8bf silu_chunk(const v4f& lo, const v4f& hi) {
  v8bf b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = (__bf16)silu_fast(lo[i]);
    b[4 + i] = (__bf16)silu_fast(hi[i]);
  }
  return b;
}

// The inference works in units of -log2(e): layer l computes
// z' = -log2(e) z (W1, b1, b2, b3 scaled by -log2(e); W2 and W3 unchanged,
// since each layer's input a' = z' sigmoid(z) = -log2(e) silu(z) carries the
// same factor), so that sigmoid(z) = 1 / (1 + 2^z') needs no multiply, and
// W4 is scaled by -ln 2 to return to the output's units.
constexpr float kNegLog2e = -1.44269504f, kNegLn2 = -0.693147181f;

__device__ __forceinline__ v8bf silu_l2_chunk(const v4f& lo, const v4f& hi) {
  v8bf b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = (__bf16)(lo[i] * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(lo[i])));
    b[4 + i] = (__bf16)(hi[i] * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(hi[i])));
  }
  return b;
}

__global__ __launch_bounds__(256) void mlp_forward_reg_kernel(const float* __restrict__ prm,
                                                              const float4* __restrict__ feat,
                                                              const float4* __restrict__ mean,
                                                              long long n,
                                                              float4* __restrict__ out) {
  __shared__ float P[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) {
    const bool w4 = i >= oW4 && i < ob4, mid = (i >= oW2 && i < ob2) || (i >= oW3 && i < ob3) || i >= ob4;
    P[i] = mid ? prm[i] : (w4 ? kNegLn2 : kNegLog2e) * prm[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  v8bf w1[4], w2[4][2], w3[4][2], w4[2];
  v4f b2[4], b3[4], b4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 8 * q + j;
      const float v = k < kIn ? P[oW1 + k * kH + 16 * t + r] : (k == kIn ? P[ob1 + 16 * t + r] : 0.f);
      w1[t][j] = (__bf16)v;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int f = kperm(c, q, j);
        w2[t][c][j] = (__bf16)P[oW2 + f * kH + 16 * t + r];
        w3[t][c][j] = (__bf16)P[oW3 + f * kH + 16 * t + r];
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      b2[t][i] = P[ob2 + 16 * t + 4 * q + i];
      b3[t][i] = P[ob3 + 16 * t + 4 * q + i];
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j)
      w4[c][j] = (__bf16)(r < kOut ? P[oW4 + kperm(c, q, j) * kOut + r] : 0.f);
#pragma unroll
  for (int i = 0; i < 4; ++i) b4[i] = (4 * q + i < kOut) ? P[ob4 + 4 * q + i] : 0.f;

  const long long groups = (n + 15) >> 4;
  const int wpb = blockDim.x >> 6;
  const long long nwaves = (long long)gridDim.x * wpb;
  // lane group 0 supplies inputs 0..7, group 1 inputs 8, 9 and the bias
  // slot 10 (= 1), groups 2 and 3 zeros; the next group's inputs (and its
  // base force) are loaded while the current group runs through the layers
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  auto fetch = [&](long long g, float4& a, float4& b, float4& c, float4& m) {
    const long long p = g * 16 + r;
    a = b = c = m = zero4;
    if (g < groups && p < n) {
      if (q == 0) {
        a = feat[3 * p];
        b = feat[3 * p + 1];
        m = mean[p];
      }
      if (q <= 1) c = feat[3 * p + 2];
    }
  };
  long long gi = (long long)blockIdx.x * wpb + (threadIdx.x >> 6);
  float4 fa, fb, fc, fm;
  fetch(gi, fa, fb, fc, fm);
  for (; gi < groups; gi += nwaves) {
    const long long p = gi * 16 + r;
    const bool live = p < n;
    float4 na, nb, nc, nm;
    fetch(gi + nwaves, na, nb, nc, nm);
    v8bf x;
    if (q == 0) {
      x[0] = (__bf16)fa.x;
      x[1] = (__bf16)fa.y;
      x[2] = (__bf16)fa.z;
      x[3] = (__bf16)fa.w;
      x[4] = (__bf16)fb.x;
      x[5] = (__bf16)fb.y;
      x[6] = (__bf16)fb.z;
      x[7] = (__bf16)fb.w;
    } else {
      x[0] = (__bf16)fc.x;
      x[1] = (__bf16)fc.y;
      x[2] = (__bf16)(q == 1 ? 1.f : 0.f);
      x[3] = (__bf16)0.f;
      x[4] = (__bf16)0.f;
      x[5] = (__bf16)0.f;
      x[6] = (__bf16)0.f;
      x[7] = (__bf16)0.f;
    }
    v4f h[4], z[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1[t], x, v4f{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    {
      const v8bf a0 = silu_l2_chunk(h[0], h[1]), a1 = silu_l2_chunk(h[2], h[3]);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][0], a0, b2[t], 0, 0, 0);
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][1], a1, z[t], 0, 0, 0);
      }
    }
    {
      const v8bf a0 = silu_l2_chunk(z[0], z[1]), a1 = silu_l2_chunk(z[2], z[3]);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][0], a0, b3[t], 0, 0, 0);
        h[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][1], a1, h[t], 0, 0, 0);
      }
    }
    v4f y;
    {
      const v8bf a0 = silu_l2_chunk(h[0], h[1]), a1 = silu_l2_chunk(h[2], h[3]);
      y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[0], a0, b4, 0, 0, 0);
      y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[1], a1, y, 0, 0, 0);
    }
    if (live && q == 0)
      out[p] = make_float4(fmaf(fc.z, y[0], fm.x), fmaf(fc.z, y[1], fm.y), fmaf(fc.z, y[2], fm.z), 0.f);
    fa = na;
    fb = nb;
    fc = nc;
    fm = nm;
  }
}

// ---------------------------------------------------------------------------
// Fused training pass: forward, loss, backward, gradient accumulation
// ---------------------------------------------------------------------------
// grad (kParams floats) and loss (1 float) are accumulated with atomics; the
// caller clears them.  wnorm = 1 / sum_p S_p^2 over the whole batch, so the
// loss is sum_p S_p^2 |y_p - t_p|^2 / sum_p S_p^2 (the squared relative RMS
// error of the predicted force when the batch is all particles).
template <bool BF16>
__global__ __launch_bounds__(256) void mlp_train_kernel(const float* __restrict__ prm,
                                                        const float4* __restrict__ feat,
                                                        const float4* __restrict__ mean,
                                                        const float4* __restrict__ target,
                                                        long long n, float wnorm,
                                                        float* __restrict__ grad,
                                                        float* __restrict__ loss) {
  extern __shared__ float sm[];
  float* X = sm;                  // TILE x kXS
  float* Z1 = X + TILE * kXS;     // TILE x 64 (pre-activations)
  float* Z2 = Z1 + TILE * kH;
  float* Z3 = Z2 + TILE * kH;
  float* DY = Z3 + TILE * kH;     // TILE x 16
  float* D1 = DY + TILE * 16;     // TILE x 64 scratch
  float* D2 = D1 + TILE * kH;     // TILE x 64 scratch
  float* G = D2 + TILE * kH;      // kParams gradient accumulator
  float* LS = G + kParams;        // 1 loss accumulator
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) G[i] = 0.f;
  if (threadIdx.x == 0) LS[0] = 0.f;
  __syncthreads();
  const long long ntile = (n + TILE - 1) / TILE;
  for (long long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long long base = tile * TILE;
    for (int i = threadIdx.x; i < TILE * kXS; i += blockDim.x) {
      const long long p = base + i / kXS;
      const int f = i % kXS;
      float v = 0.f;
      if (p < n && f < kIn) {
        const float4 a = feat[3 * p + f / 4];
        v = (&a.x)[f % 4];
      }
      X[i] = v;
    }
    __syncthreads();
    // forward
    wg_gemm<BF16>(Mat{X, kXS, 1, TILE, kIn, false}, Mat{prm + oW1, kH, 1, kIn, kH, false}, Z1, kH,
                  TILE, kH, 32, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) Z1[i] += prm[ob1 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{Z1, kH, 1, TILE, kH, true}, Mat{prm + oW2, kH, 1, kH, kH, false}, Z2, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) Z2[i] += prm[ob2 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{Z2, kH, 1, TILE, kH, true}, Mat{prm + oW3, kH, 1, kH, kH, false}, Z3, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) Z3[i] += prm[ob3 + i % kH];
    __syncthreads();
    wg_gemm<BF16>(Mat{Z3, kH, 1, TILE, kH, true}, Mat{prm + oW4, kOut, 1, kH, kOut, false}, DY, 16,
                  TILE, 16, kH, TILE, kOut, false);
    __syncthreads();
    // loss and dL/dy (zero rows beyond n)
    for (int r = threadIdx.x; r < TILE; r += blockDim.x) {
      const long long p = base + r;
      float l = 0.f;
      for (int c = 0; c < 16; ++c) {
        float d = 0.f;
        if (p < n && c < kOut) {
          const float S = feat[3 * p + 2].z;
          const float4 m = mean[p], tg = target[p];
          const float t = ((&tg.x)[c] - (&m.x)[c]) / S;
          const float w = S * S * wnorm;
          const float e = DY[r * 16 + c] + prm[ob4 + c] - t;
          l += w * e * e;
          d = 2.0f * w * e;
        }
        DY[r * 16 + c] = d;
      }
      atomicAdd(LS, l);
    }
    __syncthreads();
    // layer 4: dW4 += A3^T DY, db4 += sum DY, D1 = (DY W4^T) * silu'(Z3)
    wg_gemm<BF16>(Mat{Z3, 1, kH, kH, TILE, true}, Mat{DY, 16, 1, TILE, kOut, false}, G + oW4, kOut,
                  kH, 16, TILE, kH, kOut, true);
    wg_gemm<BF16>(Mat{DY, 16, 1, TILE, kOut, false}, Mat{prm + oW4, 1, kOut, kOut, kH, false}, D1,
                  kH, TILE, kH, 32, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < kOut; i += blockDim.x) {
      float s = 0.f;
      for (int r = 0; r < TILE; ++r) s += DY[r * 16 + i];
      G[ob4 + i] += s;
    }
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) D1[i] *= dsilu<BF16>(Z3[i]);
    __syncthreads();
    // layer 3: dW3 += A2^T D1, db3 += sum D1, D2 = (D1 W3^T) * silu'(Z2)
    wg_gemm<BF16>(Mat{Z2, 1, kH, kH, TILE, true}, Mat{D1, kH, 1, TILE, kH, false}, G + oW3, kH, kH,
                  kH, TILE, kH, kH, true);
    wg_gemm<BF16>(Mat{D1, kH, 1, TILE, kH, false}, Mat{prm + oW3, 1, kH, kH, kH, false}, D2, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
      float s = 0.f;
      for (int r = 0; r < TILE; ++r) s += D1[r * kH + i];
      G[ob3 + i] += s;
    }
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) D2[i] *= dsilu<BF16>(Z2[i]);
    __syncthreads();
    // layer 2: dW2 += A1^T D2, db2 += sum D2, D1 = (D2 W2^T) * silu'(Z1)
    wg_gemm<BF16>(Mat{Z1, 1, kH, kH, TILE, true}, Mat{D2, kH, 1, TILE, kH, false}, G + oW2, kH, kH,
                  kH, TILE, kH, kH, true);
    wg_gemm<BF16>(Mat{D2, kH, 1, TILE, kH, false}, Mat{prm + oW2, 1, kH, kH, kH, false}, D1, kH,
                  TILE, kH, kH, TILE, kH, false);
    __syncthreads();
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
      float s = 0.f;
      for (int r = 0; r < TILE; ++r) s += D2[r * kH + i];
      G[ob2 + i] += s;
    }
    for (int i = threadIdx.x; i < TILE * kH; i += blockDim.x) D1[i] *= dsilu<BF16>(Z1[i]);
    __syncthreads();
    // layer 1: dW1 += X^T D1, db1 += sum D1
    wg_gemm<BF16>(Mat{X, 1, kXS, kIn, TILE, false}, Mat{D1, kH, 1, TILE, kH, false}, G + oW1, kH, 16,
                  kH, TILE, kIn, kH, true);
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
      float s = 0.f;
      for (int r = 0; r < TILE; ++r) s += D1[r * kH + i];
      G[ob1 + i] += s;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) atomicAdd(&grad[i], G[i]);
  if (threadIdx.x == 0) atomicAdd(loss, LS[0]);
}

// ---------------------------------------------------------------------------
// Register-resident BF16 training pass (the production training kernel)
// ---------------------------------------------------------------------------
// One wave takes 16 particles per iteration through the forward of
// mlp_forward_reg_kernel (same fragments, same layout: lane (q, r) holds
// features 16t + 4q + i of particle r), then the backward in that layout:
//   dy  = 2 w (y - t)                                   (lanes q = 0)
//   d3  = (W4 dy) . silu'(z3),  d2 = (W3 d3) . silu'(z2),  d1 = (W2 d2) . silu'(z1)
// as v_mfma_f32_16x16x32_bf16 products with "transposed" weight fragments
// (A[r][k] = W[16t + r][kperm(c, q, j)]), so the deltas chain like the
// activations with no data movement.  The weight gradients sum over
// particles, which sit on the lane axis: each factor is transposed through a
// wave-private LDS tile (rows = the 16 particles, 68 bf16 per row so that
// the four lane groups read disjoint banks) and multiplied with
// v_mfma_f32_16x16x16_bf16 (K = the 16 particles) into per-wave FP32
// accumulators (160 registers).  Layer 1's bias gradient is row 10 of the
// W1 product (the input slot holding 1); the other bias gradients are
// per-lane sums of the deltas.  At the end the waves of a workgroup add
// their accumulators into one LDS copy, which is flushed to `grad` with
// global atomics.  The weight fragments live in LDS, built once per
// workgroup.
typedef short v4s __attribute__((ext_vector_type(4)));
constexpr int kFrW1 = 0, kFrW2 = 4, kFrW3 = 12, kFrW4 = 20, kFrW4T = 22, kFrW3T = 26,
              kFrW2T = 34, kFrags = 42;
constexpr int kTR = 68;                 // transpose tile row stride (bf16)
constexpr int kTile = 16 * kTR;         // one transpose tile (bf16 elements)
constexpr size_t kTrainLds = (size_t)kFrags * 64 * 16 + (size_t)9 * 64 * 16 +
                             (size_t)4 * 2 * kTile * 2 + (size_t)kParams * 4 + 16;

__device__ __forceinline__ short bf16_bits(float x) {
  return __builtin_bit_cast(short, (__bf16)x);
}

// the 16x16x16 operand of tile `tile` of a transpose tile: lane (q, r) gets
// feature 16 tile + r of particles 4q .. 4q + 3
__device__ __forceinline__ v4s tr_operand(const unsigned short* tb, int tile, int q, int r) {
  const unsigned short* b = tb + 4 * q * kTR + 16 * tile + r;
  v4s v;
  v[0] = (short)b[0];
  v[1] = (short)b[kTR];
  v[2] = (short)b[2 * kTR];
  v[3] = (short)b[3 * kTR];
  return v;
}

// four bf16 values at an 8-byte aligned tile position (row r * kTR + 4 m)
__device__ __forceinline__ void tr_put4(unsigned short* at, uint2 v) {
  *reinterpret_cast<uint2*>(at) = v;
}

// a lane's 16 features (two chunks in kperm order) into row r of a tile
__device__ __forceinline__ void tr_put_chunks(unsigned short* tb, const v8bf& c0, const v8bf& c1,
                                              int q, int r) {
  unsigned short* row = tb + r * kTR + 4 * q;
  const uint4 a = __builtin_bit_cast(uint4, c0), b = __builtin_bit_cast(uint4, c1);
  tr_put4(row, make_uint2(a.x, a.y));
  tr_put4(row + 16, make_uint2(a.z, a.w));
  tr_put4(row + 32, make_uint2(b.x, b.y));
  tr_put4(row + 48, make_uint2(b.z, b.w));
}

// orders a wave's LDS writes and reads of its transpose tiles in the
// compiler (the hardware runs one wave's LDS instructions in order)
__device__ __forceinline__ void tile_fence() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ v8bf bf16_chunk(const v4f& lo, const v4f& hi) {
  v8bf b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    b[i] = (__bf16)lo[i];
    b[4 + i] = (__bf16)hi[i];
  }
  return b;
}

__global__ __launch_bounds__(256) void mlp_train_reg_kernel(const float* __restrict__ prm,
                                                            const float4* __restrict__ feat,
                                                            const float4* __restrict__ mean,
                                                            const float4* __restrict__ target,
                                                            long long n, float wnorm,
                                                            float* __restrict__ grad,
                                                            float* __restrict__ loss) {
  extern __shared__ __align__(16) char tsm[];
  v8bf* const FR = reinterpret_cast<v8bf*>(tsm);                      // kFrags x 64 lanes
  v4f* const BI = reinterpret_cast<v4f*>(FR + kFrags * 64);           // b2[4], b3[4], b4 x 64
  unsigned short* const TB = reinterpret_cast<unsigned short*>(BI + 9 * 64);
  float* const G = reinterpret_cast<float*>(TB + 4 * 2 * kTile);       // kParams
  float* const LS = G + kParams;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, r = lane & 15;
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) G[i] = 0.f;
  if (threadIdx.x == 0) LS[0] = 0.f;
  // fragments: wave w builds those with index = w (mod 4), each lane its own
  for (int f = wave; f < kFrags; f += 4) {
    v8bf v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 8 * q + j;
      float x = 0.f;
      if (f < kFrW2) {
        const int t = f - kFrW1;
        x = k < kIn ? prm[oW1 + k * kH + 16 * t + r] : (k == kIn ? prm[ob1 + 16 * t + r] : 0.f);
      } else if (f < kFrW3) {
        const int t = (f - kFrW2) >> 1, c = (f - kFrW2) & 1;
        x = prm[oW2 + kperm(c, q, j) * kH + 16 * t + r];
      } else if (f < kFrW4) {
        const int t = (f - kFrW3) >> 1, c = (f - kFrW3) & 1;
        x = prm[oW3 + kperm(c, q, j) * kH + 16 * t + r];
      } else if (f < kFrW4T) {
        const int c = f - kFrW4;
        x = r < kOut ? prm[oW4 + kperm(c, q, j) * kOut + r] : 0.f;
      } else if (f < kFrW3T) {
        const int t = f - kFrW4T;
        x = k < kOut ? prm[oW4 + (16 * t + r) * kOut + k] : 0.f;
      } else if (f < kFrW2T) {
        const int t = (f - kFrW3T) >> 1, c = (f - kFrW3T) & 1;
        x = prm[oW3 + (16 * t + r) * kH + kperm(c, q, j)];
      } else {
        const int t = (f - kFrW2T) >> 1, c = (f - kFrW2T) & 1;
        x = prm[oW2 + (16 * t + r) * kH + kperm(c, q, j)];
      }
      v[j] = (__bf16)x;
    }
    FR[f * 64 + lane] = v;
  }
  if (wave == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      v4f a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = prm[ob2 + 16 * t + 4 * q + i];
        b[i] = prm[ob3 + 16 * t + 4 * q + i];
      }
      BI[t * 64 + lane] = a;
      BI[(4 + t) * 64 + lane] = b;
    }
    v4f c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = (4 * q + i < kOut) ? prm[ob4 + 4 * q + i] : 0.f;
    BI[8 * 64 + lane] = c;
  }
  __syncthreads();
  unsigned short* const tA = TB + wave * 2 * kTile;
  unsigned short* const tB = tA + kTile;
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  v4f gW1[4], gW2[4][4], gW3[4][4], gW4[4], gb2[4], gb3[4], gb4 = zero;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a] = gW4[a] = gb2[a] = gb3[a] = zero;
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = gW3[a][b] = zero;
  }
  float lacc = 0.f;
  const long long groups = (n + 15) >> 4;
  const long long nwaves = (long long)gridDim.x * (blockDim.x >> 6);
  // the next group's inputs are loaded while the current one is processed
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  auto fetch = [&](long long g, float4& a, float4& b, float4& c, float4& m, float4& tg) {
    const long long pp = g * 16 + r;
    a = b = c = m = tg = zero4;
    if (g < groups && pp < n) {
      if (q == 0) {
        a = feat[3 * pp];
        b = feat[3 * pp + 1];
        m = mean[pp];
        tg = target[pp];
      }
      if (q <= 1) c = feat[3 * pp + 2];
    }
  };
  long long gi = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  float4 fa, fb, fc, fm, ft;
  fetch(gi, fa, fb, fc, fm, ft);
  for (; gi < groups; gi += nwaves) {
    const long long p = gi * 16 + r;
    const bool live = p < n;
    float4 na, nb, nc, nm, nt4;
    fetch(gi + nwaves, na, nb, nc, nm, nt4);
    v8bf x;
    if (q == 0) {
      x[0] = (__bf16)fa.x;
      x[1] = (__bf16)fa.y;
      x[2] = (__bf16)fa.z;
      x[3] = (__bf16)fa.w;
      x[4] = (__bf16)fb.x;
      x[5] = (__bf16)fb.y;
      x[6] = (__bf16)fb.z;
      x[7] = (__bf16)fb.w;
    } else {
      x[0] = (__bf16)fc.x;
      x[1] = (__bf16)fc.y;
      x[2] = (__bf16)(q == 1 ? 1.f : 0.f);
#pragma unroll
      for (int j = 3; j < 8; ++j) x[j] = (__bf16)0.f;
    }
    // forward
    v4f z1[4], z2[4], z3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      z1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW1 + t) * 64 + lane], x, zero, 0, 0, 0);
    const v8bf a1c0 = silu_chunk(z1[0], z1[1]), a1c1 = silu_chunk(z1[2], z1[3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      z2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW2 + 2 * t) * 64 + lane], a1c0,
                                                      BI[t * 64 + lane], 0, 0, 0);
      z2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW2 + 2 * t + 1) * 64 + lane], a1c1,
                                                      z2[t], 0, 0, 0);
    }
    const v8bf a2c0 = silu_chunk(z2[0], z2[1]), a2c1 = silu_chunk(z2[2], z2[3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      z3[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW3 + 2 * t) * 64 + lane], a2c0,
                                                      BI[(4 + t) * 64 + lane], 0, 0, 0);
      z3[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW3 + 2 * t + 1) * 64 + lane], a2c1,
                                                      z3[t], 0, 0, 0);
    }
    const v8bf a3c0 = silu_chunk(z3[0], z3[1]), a3c1 = silu_chunk(z3[2], z3[3]);
    v4f y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[kFrW4 * 64 + lane], a3c0,
                                                    BI[8 * 64 + lane], 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW4 + 1) * 64 + lane], a3c1, y, 0, 0, 0);
    // loss and dL/dy (lanes q = 0 hold outputs 0..2 of particle r)
    v4f dy = zero;
    if (live && q == 0) {
      const float S = fc.z, w = S * S * wnorm;
      const float e0 = y[0] - (ft.x - fm.x) / S, e1 = y[1] - (ft.y - fm.y) / S,
                  e2 = y[2] - (ft.z - fm.z) / S;
      lacc += w * (e0 * e0 + e1 * e1 + e2 * e2);
      dy[0] = 2.f * w * e0;
      dy[1] = 2.f * w * e1;
      dy[2] = 2.f * w * e2;
    }
    v8bf dyb;
#pragma unroll
    for (int j = 0; j < 8; ++j) dyb[j] = (__bf16)(j < 4 ? dy[j & 3] : 0.f);
    // backward through the layers
    v4f d[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW4T + t) * 64 + lane], dyb, zero, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[t][i] *= dsilu<true>(z3[t][i]);
      gb3[t] += d[t];
    }
    const v8bf d3c0 = bf16_chunk(d[0], d[1]), d3c1 = bf16_chunk(d[2], d[3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW3T + 2 * t) * 64 + lane], d3c0, zero, 0, 0, 0);
      d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW3T + 2 * t + 1) * 64 + lane], d3c1, d[t], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[t][i] *= dsilu<true>(z2[t][i]);
      gb2[t] += d[t];
    }
    const v8bf d2c0 = bf16_chunk(d[0], d[1]), d2c1 = bf16_chunk(d[2], d[3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW2T + 2 * t) * 64 + lane], d2c0, zero, 0, 0, 0);
      d[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(FR[(kFrW2T + 2 * t + 1) * 64 + lane], d2c1, d[t], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) d[t][i] *= dsilu<true>(z1[t][i]);
    }
    const v8bf d1c0 = bf16_chunk(d[0], d[1]), d1c1 = bf16_chunk(d[2], d[3]);
    gb4 += dy;
    // weight gradients: K = the 16 particles, through the wave's two tiles
    {  // W4: a3 (64 rows) x dy (outputs 0..15, 3 used)
      tr_put_chunks(tA, a3c0, a3c1, q, r);
      const unsigned lo = (unsigned short)bf16_bits(dy[0]) | ((unsigned)(unsigned short)bf16_bits(dy[1]) << 16);
      const unsigned hi = (unsigned short)bf16_bits(dy[2]) | ((unsigned)(unsigned short)bf16_bits(dy[3]) << 16);
      tr_put4(tB + r * kTR + 4 * q, make_uint2(lo, hi));
      tile_fence();
      const v4s b = tr_operand(tB, 0, q, r);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        gW4[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(tr_operand(tA, mt, q, r), b, gW4[mt], 0, 0, 0);
      tile_fence();
    }
    {  // W3: a2 x d3
      tr_put_chunks(tA, a2c0, a2c1, q, r);
      tr_put_chunks(tB, d3c0, d3c1, q, r);
      tile_fence();
      v4s b[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) b[nt] = tr_operand(tB, nt, q, r);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const v4s a = tr_operand(tA, mt, q, r);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          gW3[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], gW3[mt][nt], 0, 0, 0);
      }
      tile_fence();
    }
    {  // W2: a1 x d2
      tr_put_chunks(tA, a1c0, a1c1, q, r);
      tr_put_chunks(tB, d2c0, d2c1, q, r);
      tile_fence();
      v4s b[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) b[nt] = tr_operand(tB, nt, q, r);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const v4s a = tr_operand(tA, mt, q, r);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          gW2[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nt], gW2[mt][nt], 0, 0, 0);
      }
      tile_fence();
    }
    {  // W1 and b1: x (inputs 0..15, slot 10 = 1) x d1
      if (q <= 1) {
        const uint4 xb = __builtin_bit_cast(uint4, x);
        tr_put4(tA + r * kTR + 8 * q, make_uint2(xb.x, xb.y));
        tr_put4(tA + r * kTR + 8 * q + 4, make_uint2(xb.z, xb.w));
      }
      tr_put_chunks(tB, d1c0, d1c1, q, r);
      tile_fence();
      const v4s a = tr_operand(tA, 0, q, r);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        gW1[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, tr_operand(tB, nt, q, r), gW1[nt], 0, 0, 0);
      tile_fence();
    }
    fa = na;
    fb = nb;
    fc = nc;
    fm = nm;
    ft = nt4;
  }
  // the waves' accumulators into the workgroup's LDS copy
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * q + i;  // within a 16-row tile
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        atomicAdd(&G[oW2 + (16 * a + row) * kH + 16 * b + r], gW2[a][b][i]);
        atomicAdd(&G[oW3 + (16 * a + row) * kH + 16 * b + r], gW3[a][b][i]);
      }
      if (row < kIn) atomicAdd(&G[oW1 + row * kH + 16 * a + r], gW1[a][i]);
      if (row == kIn) atomicAdd(&G[ob1 + 16 * a + r], gW1[a][i]);
      if (r < kOut) atomicAdd(&G[oW4 + (16 * a + row) * kOut + r], gW4[a][i]);
      atomicAdd(&G[ob2 + 16 * a + row], gb2[a][i]);
      atomicAdd(&G[ob3 + 16 * a + row], gb3[a][i]);
    }
  }
  if (q == 0) {
#pragma unroll
    for (int i = 0; i < kOut; ++i) atomicAdd(&G[ob4 + i], gb4[i]);
  }
  atomicAdd(LS, lacc);
  __syncthreads();
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) atomicAdd(&grad[i], G[i]);
  if (threadIdx.x == 0) atomicAdd(loss, LS[0]);
}

// Adam on fp32 parameters (bias-corrected).
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ prm,
                                                   const float* __restrict__ grad,
                                                   float* __restrict__ m1, float* __restrict__ m2,
                                                   int np, float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float g = grad[i];
  const float a = b1 * m1[i] + (1.f - b1) * g;
  const float b = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = a;
  m2[i] = b;
  prm[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// FP64 direct summation of the softened Newtonian potential gradient at a
// subsample of target particles, minimum-image in the periodic box:
// grad_i = G_eff sum_j m_j d_ij / (|d_ij|^2 + eps^2)^(3/2), d_ij = r_i - r_j
// wrapped into [-W/2, W/2).  One workgroup per target; sources are streamed
// by the workgroup's threads and reduced in LDS.
// FP64 direct summation at a few target particles.  A workgroup takes
// kDirectTargets targets (their positions in FP64 registers) and streams all
// n sources once, so the source array is read n_targets / kDirectTargets
// times from cache instead of n_targets times.  Per pair: minimum-image
// separation d (d -= W floor((d + W/2) / W), with the division taken as a
// multiplication by the FP64 reciprocal), r2 = |d|^2 + eps2, and
// r2^(-3/2) from the float32 reciprocal square root refined by two
// Newton-Raphson steps in FP64 (relative error below 1e-15, far under the
// audit's needs); the self term is skipped.  Sums are fixed-order.
constexpr int kDirectTargets = 8;

__global__ __launch_bounds__(256) void direct_sum_kernel(const float4* __restrict__ pos,
                                                         long long n,
                                                         const int* __restrict__ targets, int nt,
                                                         double W, double G_eff, double eps2,
                                                         double* __restrict__ out) {
  __shared__ double red[256];
  const int t0 = blockIdx.x * kDirectTargets;
  double tx[kDirectTargets], ty[kDirectTargets], tz[kDirectTargets];
  long long ti[kDirectTargets];
  bool valid[kDirectTargets];
#pragma unroll
  for (int i = 0; i < kDirectTargets; ++i) {
    const int t = min(t0 + i, nt - 1);
    ti[i] = targets[t];
    valid[i] = t0 + i < nt && ti[i] >= 0 && ti[i] < n;
    const float4 p = valid[i] ? pos[ti[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
    tx[i] = p.x;
    ty[i] = p.y;
    tz[i] = p.z;
  }
  double acc[kDirectTargets][3];
#pragma unroll
  for (int i = 0; i < kDirectTargets; ++i) acc[i][0] = acc[i][1] = acc[i][2] = 0.0;
  const double inv_w = 1.0 / W;
  for (long long j = threadIdx.x; j < n; j += blockDim.x) {
    const float4 pj = pos[j];
    const double sx = pj.x, sy = pj.y, sz = pj.z, m = pj.w;
#pragma unroll
    for (int i = 0; i < kDirectTargets; ++i) {
      double dx = tx[i] - sx, dy = ty[i] - sy, dz = tz[i] - sz;
      dx = fma(-W, floor(fma(dx, inv_w, 0.5)), dx);
      dy = fma(-W, floor(fma(dy, inv_w, 0.5)), dy);
      dz = fma(-W, floor(fma(dz, inv_w, 0.5)), dz);
      const bool self = j == ti[i];
      double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      r2 = self ? 1.0 : r2;
      double y = (double)__frsqrt_rn((float)r2);
      y = y * fma(-0.5 * r2, y * y, 1.5);
      y = y * fma(-0.5 * r2, y * y, 1.5);
      if (!(r2 > 1e-30 && r2 < 1e30)) y = 1.0 / sqrt(r2);  // outside float32's normal range
      const double inv3 = self ? 0.0 : m * (y * y * y);
      acc[i][0] = fma(dx, inv3, acc[i][0]);
      acc[i][1] = fma(dy, inv3, acc[i][1]);
      acc[i][2] = fma(dz, inv3, acc[i][2]);
    }
  }
#pragma unroll
  for (int i = 0; i < kDirectTargets; ++i) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      red[threadIdx.x] = acc[i][d];
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0 && t0 + i < nt)
        out[3 * (t0 + i) + d] = valid[i] ? G_eff * red[0] : __builtin_nan("");
      __syncthreads();
    }
  }
}

hipError_t launch_direct_sum(const float4* pos, long long n, const int* targets, int nt, double W,
                             double G_eff, double eps2, double* out, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  hipLaunchKernelGGL(direct_sum_kernel, dim3((unsigned)((nt + kDirectTargets - 1) / kDirectTargets)),
                     dim3(256), 0, s, pos, n, targets, nt, W, G_eff, eps2, out);
  return hipGetLastError();
}

hipError_t launch_closure_features(const Params& P, const float4* pos, long long n,
                                   const float* rot, const float* off, const float* frot,
                                   const float4* force, const float4* force_dc,
                                   const long long* grids, float rho0, const int* order,
                                   float4* fm, float4* blk, float4* feat, float4* mean,
                                   hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long total = (long long)P.k * P.cells;
  hipLaunchKernelGGL(closure_grids_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     force, force_dc, grids, total, P.lg, fm, blk);
  const dim3 grid((unsigned)((n + 255) / 256));
  if ((P.shift_ok & kShiftF4) && (P.shift_ok & kShiftBlk))
    hipLaunchKernelGGL(closure_features_kernel<true>, grid, dim3(256), 0, s, pos, n, rot, off, frot,
                       fm, blk, P, rho0, order, feat, mean);
  else
    hipLaunchKernelGGL(closure_features_kernel<false>, grid, dim3(256), 0, s, pos, n, rot, off,
                       frot, fm, blk, P, rho0, order, feat, mean);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* prm, const float4* feat, const float4* mean,
                              long long n, float4* out, bool bf16, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (bf16) {
    // 4 waves per workgroup, 16 particles per wave iteration; the kernel
    // holds ~230 registers per lane (2 waves per SIMD), so 512 workgroups
    // (2 per CU, 36 KB of LDS each for the staged parameters) fill the chip.
    const long long groups = (n + 15) / 16;
    const long long want = (groups + 3) / 4;
    const unsigned blocks = (unsigned)(want < 512 ? want : 512);
    hipLaunchKernelGGL(mlp_forward_reg_kernel, dim3(blocks), dim3(256), 0, s, prm, feat, mean, n,
                       out);
  } else {
    const long long tiles = (n + TILE - 1) / TILE;
    const unsigned blocks = (unsigned)(tiles < 4096 ? tiles : 4096);
    hipLaunchKernelGGL(mlp_forward_kernel<false>, dim3(blocks), dim3(256), 0, s, prm, feat, mean, n,
                       out);
  }
  return hipGetLastError();
}

size_t mlp_param_count() { return (size_t)kParams; }

hipError_t launch_mlp_train(const float* prm, const float4* feat, const float4* mean,
                            const float4* target, long long n, float wnorm, float* grad,
                            float* loss, bool bf16, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (bf16) {
    // one workgroup of 4 waves per CU (the kernel holds its gradient
    // accumulators in registers: one wave per SIMD), 16 particles per wave
    // iteration
    hipError_t e = hipFuncSetAttribute((const void*)mlp_train_reg_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrainLds);
    if (e != hipSuccess) return e;
    const long long groups = (n + 15) / 16, want = (groups + 3) / 4;
    const unsigned blocks = (unsigned)(want < 256 ? want : 256);
    hipLaunchKernelGGL(mlp_train_reg_kernel, dim3(blocks), dim3(256), kTrainLds, s, prm, feat, mean,
                       target, n, wnorm, grad, loss);
    return hipGetLastError();
  }
  const size_t lds = (size_t)(TILE * kXS + 3 * TILE * kH + TILE * 16 + 2 * TILE * kH + kParams + 1) *
                     sizeof(float);
  hipError_t e = hipFuncSetAttribute((const void*)mlp_train_kernel<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  if (e != hipSuccess) return e;
  const long long tiles = (n + TILE - 1) / TILE;
  const unsigned blocks = (unsigned)(tiles < 256 ? tiles : 256);
  hipLaunchKernelGGL(mlp_train_kernel<false>, dim3(blocks), dim3(256), lds, s, prm, feat, mean,
                     target, n, wnorm, grad, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float* prm, const float* grad, float* m1, float* m2, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, prm, grad, m1, m2,
                     kParams, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace mog
