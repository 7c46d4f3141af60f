// Spectral Poisson solve on MFMA (g >= 16).
//
// Same three passes as spectral.hip (A: x- and y-DFT of a z-slab; B: z-DFT,
// Green's function x i*k, inverse z-DFT of a y-slab; C: inverse y- and
// x-DFT of a z-slab and the float4 pack), but every length-g transform is a
// dense complex matrix product Out = W . In on the matrix cores:
//   Re(Out) = [Re W | -Im W] . [Re In ; Im In]
//   Im(Out) = [Im W |  Re W] . [Re In ; Im In]
// i.e. two real GEMMs of K = 2g (K = g for real input) per 16x16 output
// tile, the DFT matrix as the A operand (FP64-derived table) and the data,
// held in LDS as padded real/imaginary planes, as the B operand.  The host
// also stores the A operand in each lane's register layout after the FP32
// row tables: split (BF16 hi/lo) or converted (FP16) for the 16-bit
// products, reordered for the FP32 ones, so a lane's operand is one
// coalesced load (16 bytes per k-block, 4 bytes per k-step) instead of
// scattered row reads and their conversions per tile.  The DFT
// matrix is symmetric, so the transform along the fast axis x uses the same
// product on the transposed slab (strides swapped on read and write).
//
// Precisions of the products (accumulation is FP32 throughout):
//   FP32    v_mfma_f32_16x16x4_f32
//   BF16X3  v_mfma_f32_16x16x32_bf16, a = a_hi + a_lo split, three products
//           hi*hi + hi*lo + lo*hi (the lo*lo term is dropped)
//   FP16    v_mfma_f32_16x16x32_f16, one product; each slab is scaled by a
//           power of two so its largest magnitude sits near 2^14
// MFMA lane layouts (checked on gfx950): 16x16x4 f32 A[l%16][l/16],
// B[l/16][l%16]; 16x16x32 A[l%16][8(l/16)+j], B[8(l/16)+j][l%16]; the
// accumulator holds C[4(l/16)+i][l%16], i < 4.
#include "mog_common.h"

namespace mog {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

enum { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP16 = 2 };

// An LDS plane pair: element (r, c) of the slab matrix is re[r*rs + c*cs]
// and re[r*rs + c*cs + im].
struct Planes {
  float* p;
  int rs, cs, im;
  __device__ float& re(int r, int c) const { return p[r * rs + c * cs]; }
  __device__ float& imv(int r, int c) const { return p[r * rs + c * cs + im]; }
};

// Element k (< 2g) of column c of [Re In ; Im In] (zero when the input is
// real and k >= g).
__device__ __forceinline__ float bval(const Planes& in, int g, bool real_in, int k, int c) {
  if (k < g) return in.re(k, c);
  return real_in ? 0.0f : in.imv(k - g, c);
}

// One wave computes the output tiles t = t0, t0 + tstep, ... of
// Out = W . In.  W points at a direction's table: the two K = 2g row tables
// WA = [Re W | -Im W] and WB = [Im W | Re W] (row-major, g rows of 2g
// floats), then their A-operand fragments, which are what the products read.
// `scale` multiplies the data before FP16 conversion (1 otherwise); the
// result is multiplied by 1/scale.
template <int PREC>
__device__ void wave_cgemm(const float* __restrict__ W, int g, const Planes in, bool real_in,
                           const Planes out, int t0, int tstep, float scale) {
  const int lane = threadIdx.x & 63;
  const int tpr = g >> 4;  // tiles per row of tiles
  const int nt = tpr * tpr;
  const int K = real_in ? g : 2 * g;
  for (int t = t0; t < nt; t += tstep) {
    const int r0 = (t / tpr) * 16, c0 = (t % tpr) * 16;
    v4f ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
    const int bcol = c0 + (lane & 15);
    if (PREC == PREC_FP32) {
      // A operands from the host-built FP32 fragments (after the 16-bit ones)
      const float* AF32 = W + 4 * g * g + 6 * tpr * tpr * 256;
      const int nk = g >> 1, nf32 = tpr * nk * 64;
      const int rt = r0 >> 4;
      for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + (lane >> 4);
        const int fi = ((rt * nk + (k0 >> 2)) << 6) + lane;
        const float a1 = AF32[fi];
        const float a2 = AF32[nf32 + fi];
        const float b = bval(in, g, real_in, k, bcol);
        ar = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, ar, 0, 0, 0);
        ai = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b, ai, 0, 0, 0);
      }
    } else {
      // A operands: the host-built fragments that follow the FP32 tables
      const uint4* AF = reinterpret_cast<const uint4*>(W + 4 * g * g);
      const int nf = tpr * tpr * 64;  // fragments per part
      const int rt = r0 >> 4;
      for (int k0 = 0; k0 < K; k0 += 32) {
        const int kb = k0 + 8 * (lane >> 4);
        const int fi = ((rt * tpr + (k0 >> 5)) << 6) + lane;
        float fb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = kb + j;
          fb[j] = k < K ? bval(in, g, real_in, k, bcol) * scale : 0.0f;
        }
        if (PREC == PREC_BF16X3) {
          const v8bf a1h = __builtin_bit_cast(v8bf, AF[fi]);
          const v8bf a1l = __builtin_bit_cast(v8bf, AF[nf + fi]);
          const v8bf a2h = __builtin_bit_cast(v8bf, AF[2 * nf + fi]);
          const v8bf a2l = __builtin_bit_cast(v8bf, AF[3 * nf + fi]);
          v8bf bh, bl;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            bh[j] = (__bf16)fb[j];
            bl[j] = (__bf16)(fb[j] - (float)bh[j]);
          }
          ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1l, bh, ar, 0, 0, 0);
          ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, bl, ar, 0, 0, 0);
          ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1h, bh, ar, 0, 0, 0);
          ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2l, bh, ai, 0, 0, 0);
          ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h, bl, ai, 0, 0, 0);
          ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h, bh, ai, 0, 0, 0);
        } else {
          const v8h a1 = __builtin_bit_cast(v8h, AF[4 * nf + fi]);
          const v8h a2 = __builtin_bit_cast(v8h, AF[5 * nf + fi]);
          v8h b;
#pragma unroll
          for (int j = 0; j < 8; ++j) b[j] = (_Float16)fb[j];
          ar = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b, ar, 0, 0, 0);
          ai = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2, b, ai, 0, 0, 0);
        }
      }
    }
    const float inv = 1.0f / scale;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 4 * (lane >> 4) + i;
      out.re(r, bcol) = ar[i] * inv;
      out.imv(r, bcol) = ai[i] * inv;
    }
  }
}

// Power-of-two scale that maps the largest magnitude of a g x g slab
// (padded row stride ld; imaginary plane at +im when complex) near 2^14
// (FP16 only; 1 otherwise).  Called by the whole workgroup.
template <int PREC>
__device__ float slab_scale(const float* p, int g, int lg, int ld, int im, bool complex_in,
                            float* red) {
  if (PREC != PREC_FP16) return 1.0f;
  float m = 0.0f;
  for (int i = threadIdx.x; i < g * g; i += blockDim.x) {
    const int at = (i >> lg) * ld + (i & (g - 1));
    m = fmaxf(m, fabsf(p[at]));
    if (complex_in) m = fmaxf(m, fabsf(p[at + im]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  float mm = 0.0f;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mm = fmaxf(mm, red[w]);
  if (!(mm > 0.0f) || !isfinite(mm)) return 1.0f;
  int e;
  frexpf(mm, &e);  // mm = f * 2^e, f in [0.5, 1)
  return ldexpf(1.0f, 14 - e);
}

// Pass A: one workgroup per (grid, z) slab.  LDS: X (real, padded g x g+1),
// T (complex), S (complex), reduction scratch.
template <int PREC>
__global__ __launch_bounds__(256) void mfma_fwd_xy_kernel(const long long* __restrict__ mass,
                                                          float2* __restrict__ S,
                                                          const float* __restrict__ Wf, int g,
                                                          int lg) {
  extern __shared__ float sm[];
  const int ld = g + 1, plane = g * ld;
  float* X = sm;
  float* T = X + plane;      // re, im at +plane
  float* U = T + 2 * plane;  // re, im at +plane
  float* red = U + 2 * plane;
  const float* W = Wf;
  const int gg = g * g, m = g - 1;
  const long long base = (long long)blockIdx.x * gg;
  for (int i = threadIdx.x; i < gg; i += blockDim.x)
    X[(i >> lg) * ld + (i & m)] = (float)mass[base + i];
  __syncthreads();
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  float sc = slab_scale<PREC>(X, g, lg, ld, 0, false, red);
  // x: T = X . W  <=>  T^T = W . X^T
  wave_cgemm<PREC>(W, g, Planes{X, 1, ld, 0}, true, Planes{T, 1, ld, plane}, wave, nw, sc);
  __syncthreads();
  sc = slab_scale<PREC>(T, g, lg, ld, plane, true, red);
  // y: U = W . T
  wave_cgemm<PREC>(W, g, Planes{T, ld, 1, plane}, false, Planes{U, ld, 1, plane}, wave, nw,
                   sc);
  __syncthreads();
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int r = i >> lg, c = i & m;
    S[base + i] = make_float2(U[r * ld + c], U[r * ld + c + plane]);
  }
}

// Pass B: one workgroup per (grid, ky): z-DFT, Green x i*k, inverse z-DFT.
template <int PREC>
__global__ __launch_bounds__(256) void mfma_z_green_kernel(float2* __restrict__ S,
                                                           float2* __restrict__ S2,
                                                           const float* __restrict__ Wf,
                                                           const float* __restrict__ Wi,
                                                           const float* __restrict__ gtab,
                                                           const float* __restrict__ kd,
                                                           const float2* __restrict__ conv,
                                                           int g, int lg, int need_x) {
  extern __shared__ float sm[];
  const int ld = g + 1, plane = g * ld;
  float* L = sm;              // [z][kx]
  float* R = L + 2 * plane;   // [kz][kx]
  float* Z = R + 2 * plane;   // y + i z spectrum
  float* X = Z + 2 * plane;   // x spectrum
  float* red = X + 2 * plane;
  const int gg = g * g, m = g - 1;
  const int grid = blockIdx.x >> lg;
  const int ky = blockIdx.x & m;
  const long long gbase = (long long)grid * g * gg;
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int z = i >> lg, kx = i & m;
    const float2 v = S[gbase + (long long)z * gg + ky * g + kx];
    L[z * ld + kx] = v.x;
    L[z * ld + kx + plane] = v.y;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  float sc = slab_scale<PREC>(L, g, lg, ld, plane, true, red);
  wave_cgemm<PREC>(Wf, g, Planes{L, ld, 1, plane}, false, Planes{R, ld, 1, plane},
                   wave, nw, sc);
  __syncthreads();
  const float kyv = kd[ky];
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int kz = i >> lg, kx = i & m;
    const int fi = (kz * g + ky) * g + kx;
    if (conv) {
      // real-space kernel spectrum: Z = rho * (T_y + i T_z), X = rho * T_x
      const float ar = R[kz * ld + kx], ai = R[kz * ld + kx + plane];
      const float2 t = conv[fi], u = conv[g * gg + fi];
      Z[kz * ld + kx] = fmaf(ar, t.x, -(ai * t.y));
      Z[kz * ld + kx + plane] = fmaf(ar, t.y, ai * t.x);
      X[kz * ld + kx] = fmaf(ar, u.x, -(ai * u.y));
      X[kz * ld + kx + plane] = fmaf(ar, u.y, ai * u.x);
      continue;
    }
    const float gf = gtab[fi];
    const float pr = R[kz * ld + kx] * gf, pi = R[kz * ld + kx + plane] * gf;
    const float kzv = kd[kz];
    Z[kz * ld + kx] = fmaf(-pr, kzv, -(pi * kyv));
    Z[kz * ld + kx + plane] = fmaf(pr, kyv, -(pi * kzv));
    if (need_x) {
      const float kxv = kd[kx];
      X[kz * ld + kx] = -(pi * kxv);
      X[kz * ld + kx + plane] = pr * kxv;
    }
  }
  __syncthreads();
  sc = slab_scale<PREC>(Z, g, lg, ld, plane, true, red);
  wave_cgemm<PREC>(Wi, g, Planes{Z, ld, 1, plane}, false, Planes{R, ld, 1, plane},
                   wave, nw, sc);
  if (need_x) {
    __syncthreads();
    sc = slab_scale<PREC>(X, g, lg, ld, plane, true, red);
    wave_cgemm<PREC>(Wi, g, Planes{X, ld, 1, plane}, false,
                     Planes{L, ld, 1, plane}, wave, nw, sc);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int z = i >> lg, kx = i & m;
    const long long at = gbase + (long long)z * gg + ky * g + kx;
    S[at] = make_float2(R[z * ld + kx], R[z * ld + kx + plane]);
    if (need_x) S2[at] = make_float2(L[z * ld + kx], L[z * ld + kx + plane]);
  }
}

// Pass C: one workgroup per (grid, z): inverse y, inverse x, pack float4
// (rotated to the world frame when `world`).
template <int PREC>
__global__ __launch_bounds__(256) void mfma_inv_xy_kernel(const float2* __restrict__ S,
                                                          const float2* __restrict__ S2,
                                                          float4* __restrict__ force,
                                                          const float* __restrict__ Wi, int g,
                                                          int lg, int need_x,
                                                          const float* __restrict__ rot,
                                                          int world) {
  extern __shared__ float sm[];
  const int ld = g + 1, plane = g * ld;
  float* Z = sm;
  float* T = Z + 2 * plane;
  float* X = T + 2 * plane;
  float* red = X + 2 * plane;
  const float* W = Wi;
  const int gg = g * g, m = g - 1;
  const long long base = (long long)blockIdx.x * gg;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // y + i z field: inverse y then inverse x; result back in Z
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int r = i >> lg, c = i & m;
    const float2 v = S[base + i];
    Z[r * ld + c] = v.x;
    Z[r * ld + c + plane] = v.y;
  }
  __syncthreads();
  float sc = slab_scale<PREC>(Z, g, lg, ld, plane, true, red);
  wave_cgemm<PREC>(W, g, Planes{Z, ld, 1, plane}, false, Planes{T, ld, 1, plane}, wave, nw, sc);
  __syncthreads();
  sc = slab_scale<PREC>(T, g, lg, ld, plane, true, red);
  wave_cgemm<PREC>(W, g, Planes{T, 1, ld, plane}, false, Planes{Z, 1, ld, plane}, wave, nw, sc);
  if (need_x) {
    __syncthreads();
    for (int i = threadIdx.x; i < gg; i += blockDim.x) {
      const int r = i >> lg, c = i & m;
      const float2 v = S2[base + i];
      X[r * ld + c] = v.x;
      X[r * ld + c + plane] = v.y;
    }
    __syncthreads();
    sc = slab_scale<PREC>(X, g, lg, ld, plane, true, red);
    wave_cgemm<PREC>(W, g, Planes{X, ld, 1, plane}, false, Planes{T, ld, 1, plane}, wave, nw,
                     sc);
    __syncthreads();
    sc = slab_scale<PREC>(T, g, lg, ld, plane, true, red);
    wave_cgemm<PREC>(W, g, Planes{T, 1, ld, plane}, false, Planes{X, 1, ld, plane}, wave, nw,
                     sc);
  }
  __syncthreads();
  const float* Rm = rot + 12 * (blockIdx.x >> lg);
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int r = i >> lg, c = i & m;
    const float fy = Z[r * ld + c], fz = Z[r * ld + c + plane];
    const float fx = need_x ? X[r * ld + c] : fz;
    const float mag = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
    if (world) {
      const float wx = fmaf(Rm[0], fx, fmaf(Rm[1], fy, Rm[2] * fz));
      const float wy = fmaf(Rm[4], fx, fmaf(Rm[5], fy, Rm[6] * fz));
      const float wz = fmaf(Rm[8], fx, fmaf(Rm[9], fy, Rm[10] * fz));
      force[base + i] = make_float4(wx, wy, wz, mag);
    } else {
      force[base + i] = make_float4(fx, fy, fz, mag);
    }
  }
}

template <int PREC>
static hipError_t launch_mfma_prec(const Params& P, const long long* grids, float2* S1,
                                   float2* S2, float4* force, const float* Wf, const float* Wi,
                                   const float* gtab, const float* kd, const float2* conv,
                                   int need_x, const float* rot, int world, hipStream_t s) {
  const int g = P.g, lg = P.lg;
  const unsigned slabs = (unsigned)(P.k * g);
  const int threads = g == 16 ? 64 : 256;
  const size_t plane = (size_t)g * (g + 1) * sizeof(float);
  const size_t red = 64 * sizeof(float);
  const size_t ldsA = 5 * plane + red, ldsB = 8 * plane + red, ldsC = 6 * plane + red;
  hipError_t e;
  if ((e = hipFuncSetAttribute((const void*)mfma_fwd_xy_kernel<PREC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)mfma_z_green_kernel<PREC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)mfma_inv_xy_kernel<PREC>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(mfma_fwd_xy_kernel<PREC>, dim3(slabs), dim3(threads), ldsA, s, grids, S1, Wf,
                     g, lg);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(mfma_z_green_kernel<PREC>, dim3(slabs), dim3(threads), ldsB, s, S1, S2, Wf, Wi,
                     gtab, kd, conv, g, lg, need_x);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(mfma_inv_xy_kernel<PREC>, dim3(slabs), dim3(threads), ldsC, s, S1, S2, force,
                     Wi, g, lg, need_x, rot, world);
  return hipGetLastError();
}

hipError_t launch_solve_mfma(const Params& P, int precision, const long long* grids, float2* S1,
                             float2* S2, float4* force, const float* Wf, const float* Wi,
                             const float* gtab, const float* kd, const float2* conv, int need_x,
                             const float* rot, int world, hipStream_t s) {
  switch (precision) {
    case PREC_BF16X3:
      return launch_mfma_prec<PREC_BF16X3>(P, grids, S1, S2, force, Wf, Wi, gtab, kd, conv, need_x, rot,
                                           world, s);
    case PREC_FP16:
      return launch_mfma_prec<PREC_FP16>(P, grids, S1, S2, force, Wf, Wi, gtab, kd, conv, need_x, rot,
                                         world, s);
    default:
      return launch_mfma_prec<PREC_FP32>(P, grids, S1, S2, force, Wf, Wi, gtab, kd, conv, need_x, rot,
                                         world, s);
  }
}

}  // namespace mog
