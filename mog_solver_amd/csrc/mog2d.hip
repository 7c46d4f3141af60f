// 2D MOG force (the DIM == 2 branches of update.glsl:19-34, pack.glsl:8 and
// init_particles.glsl:151, which the reference never compiles: main.py:237
// always defines DIM 3).  k planar grids of g x g cells over [0, W)^2:
//
//   deposit  NGP int64 milli-mass, one workgroup per (particle chunk, grid)
//            with the g^2 histogram in LDS (u32 partials, 64-bit flush);
//            cells from the fast path of the 3D kernels (two FMAs per
//            axis, face test, exact fallback);
//   solve    one workgroup per grid, the whole grid in LDS: real x-DFT,
//            y-DFT, Green's function -2 pi G / k^2 x i k (Nyquist-positive
//            wave numbers), one inverse transform of Fx + i Fy (the
//            derivative factor's Nyquist entry zeroed: that row only feeds
//            the imaginary parts Re() discards), 1/(1000 g^2) folded in;
//   gather   one thread per particle over all grids: same cell as the
//            deposit, world force M F, average over k.
//
// Cell index (exact float32 sequence, no FMA): c = p.xy - W/2;
// r_m = rot[m] c0 + rot[4+m] c1 (the 2x2 block of the packed 3D layout,
// mat2(col0.xy, col1.xy) * c); l = (r + W/2) - off; per axis
// int(mod(mod(l, W) + W, W) / h) % g.
#include "mog_common.h"

namespace mog {

__device__ __forceinline__ int cell2d(float c0, float c1, const float* __restrict__ R,
                                      const float* __restrict__ O, float W, float half, float h,
                                      int lg) {
  float r0 = R[0] * c0;
  r0 = r0 + R[4] * c1;
  float r1 = R[1] * c0;
  r1 = r1 + R[5] * c1;
  const float l0 = (r0 + half) - O[0];
  const float l1 = (r1 + half) - O[1];
  const int gm = (1 << lg) - 1;
  return (axis_cell_exact(l1, W, h, gm) << lg) | axis_cell_exact(l0, W, h, gm);
}

// Fast 2D cell from the first two rows of the grid's fast row (the 3D row of
// make_fast_row; its z coefficients are zero in the planar layout), with
// the 3D path's face test: w_m = l_m / h - 1/2 as two FMAs, cell floor(v) =
// rint(w).  Two products and one add round less than the 3D sequence's
// three and two, so the 3D margin bounds this path too.  Falls back to the
// exact sequence for undecided pairs and particles outside |c| <= CB.
__device__ __forceinline__ int cell2d_index(float c0, float c1, const float* __restrict__ F,
                                            const float* __restrict__ R,
                                            const float* __restrict__ O, float W, float half,
                                            float h, int lg, float lim, float CB) {
  const float w0 = fmaf(F[1], c1, fmaf(F[0], c0, F[3]));
  const float w1 = fmaf(F[5], c1, fmaf(F[4], c0, F[7]));
  const float r0 = rintf(w0), r1 = rintf(w1);
  const bool ok = (fmaxf(fabsf(w0 - r0), fabsf(w1 - r1)) < lim) & (fabsf(c0) <= CB) &
                  (fabsf(c1) <= CB);
  const int gm = (1 << lg) - 1;
  if (ok) return (((int)r1 & gm) << lg) | ((int)r0 & gm);
  return cell2d(c0, c1, R, O, W, half, h, lg);
}

__global__ __launch_bounds__(256) void deposit2d_kernel(const float4* __restrict__ pos,
                                                        long long n, long long chunk,
                                                        const float* __restrict__ rot,
                                                        const float* __restrict__ off,
                                                        const float* __restrict__ frot, float W,
                                                        float h, int lg, float lim,
                                                        unsigned long long* __restrict__ grids) {
  extern __shared__ unsigned int hist[];
  const int cells = 1 << (2 * lg);
  const int grid = blockIdx.y;
  for (int i = threadIdx.x; i < cells; i += blockDim.x) hist[i] = 0u;
  __syncthreads();
  const float half = W * 0.5f;
  const float* R = rot + 12 * grid;
  const float* O = off + 4 * grid;
  unsigned long long* dst = grids + (long long)grid * cells;
  const long long p0 = (long long)blockIdx.x * chunk;
  const long long p1 = min(n, p0 + chunk);
  const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)(p1 - p0);
  for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const float4 q = pos[p];
    const unsigned long long milli = milli_of(q.w);
    if (milli == 0ull) continue;
    const int cell =
        cell2d_index(q.x - half, q.y - half, frot + 12 * grid, R, O, W, half, h, lg, lim, 1.5f * W);
    if (milli <= lds_limit)
      atomicAdd(&hist[cell], (unsigned int)milli);
    else
      atomicAdd(&dst[cell], milli);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < cells; i += blockDim.x) {
    const unsigned int v = hist[i];
    if (v) atomicAdd(&dst[i], (unsigned long long)v);
  }
}

// One workgroup per grid.  tw[j] = (cos, sin)(2 pi j / g); gtab[(ky g + kx)]
// = -2 pi G / k^2 / (1000 g^2) (0 at DC); kd[i] = Nyquist-positive wave
// number with the Nyquist entry zeroed.  Output (Fx, Fy) per cell.
__global__ __launch_bounds__(256) void solve2d_kernel(const long long* __restrict__ grids,
                                                      const float2* __restrict__ tw,
                                                      const float* __restrict__ gtab,
                                                      const float* __restrict__ kd, int lg,
                                                      float2* __restrict__ force) {
  extern __shared__ float2 sm2[];
  const int g = 1 << lg, gg = g * g, m = g - 1;
  float2* twl = sm2;          // g
  float* re = (float*)(sm2 + g);  // g*g reals (as g*g/2 float2 slots)
  float2* A = sm2 + g + gg / 2 + 1;
  float2* B = A + gg;
  const long long base = (long long)blockIdx.x * gg;
  for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
  for (int i = threadIdx.x; i < gg; i += blockDim.x) re[i] = (float)grids[base + i];
  __syncthreads();
  // x: A[y][kx] = sum_x re[y][x] e^{-2 pi i x kx / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, kx = o & m;
    float ar = 0.f, ai = 0.f;
    for (int x = 0; x < g; ++x) {
      const float a = re[y * g + x];
      const float2 w = twl[(x * kx) & m];
      ar = fmaf(a, w.x, ar);
      ai = fmaf(-a, w.y, ai);
    }
    A[o] = make_float2(ar, ai);
  }
  __syncthreads();
  // y + Green + derivative: B[ky][kx] = phi (-ky' + i kx') with
  // phi = gtab * sum_y A[y][kx] e^{-2 pi i y ky / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int ky = o >> lg, kx = o & m;
    float sr = 0.f, si = 0.f;
    for (int y = 0; y < g; ++y) {
      const float2 a = A[y * g + kx];
      const float2 w = twl[(y * ky) & m];
      sr += fmaf(a.x, w.x, a.y * w.y);
      si += fmaf(a.y, w.x, -(a.x * w.y));
    }
    const float gf = gtab[o];
    const float pr = sr * gf, pi = si * gf;
    const float kxv = kd[kx], kyv = kd[ky];
    // i kx phi + i (i ky phi) = phi * (-ky + i kx)
    B[o] = make_float2(fmaf(-pr, kyv, -(pi * kxv)), fmaf(pr, kxv, -(pi * kyv)));
  }
  __syncthreads();
  // inverse y: A[y][kx] = sum_ky B[ky][kx] e^{+2 pi i y ky / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, kx = o & m;
    float sr = 0.f, si = 0.f;
    for (int ky = 0; ky < g; ++ky) {
      const float2 b = B[ky * g + kx];
      const float2 w = twl[(y * ky) & m];
      sr += fmaf(b.x, w.x, -(b.y * w.y));
      si += fmaf(b.y, w.x, b.x * w.y);
    }
    A[o] = make_float2(sr, si);
  }
  __syncthreads();
  // inverse x: (Fx, Fy) = sum_kx A[y][kx] e^{+2 pi i x kx / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, x = o & m;
    float fx = 0.f, fy = 0.f;
    for (int kx = 0; kx < g; ++kx) {
      const float2 a = A[y * g + kx];
      const float2 w = twl[(x * kx) & m];
      fx += fmaf(a.x, w.x, -(a.y * w.y));
      fy += fmaf(a.y, w.x, a.x * w.y);
    }
    force[base + o] = make_float2(fx, fy);
  }
}

__global__ __launch_bounds__(256) void gather2d_kernel(const float4* __restrict__ pos,
                                                       long long n, int k,
                                                       const float* __restrict__ rot,
                                                       const float* __restrict__ off,
                                                       const float* __restrict__ frot, float W,
                                                       float h, int lg, float lim,
                                                       const float2* __restrict__ force,
                                                       float4* __restrict__ avg) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 q = pos[i];
  const float half = W * 0.5f;
  const float c0 = q.x - half, c1 = q.y - half;
  const long long cells = 1ll << (2 * lg);
  float tx = 0.f, ty = 0.f;
  for (int gr = 0; gr < k; ++gr) {
    const float* R = rot + 12 * gr;
    const int cell = cell2d_index(c0, c1, frot + 12 * gr, R, off + 4 * gr, W, half, h, lg, lim,
                                  1.5f * W);
    const float2 F = force[gr * cells + cell];
    // transpose(rot) * F: world_m = sum_j rot[4m + j] F_j
    tx += fmaf(R[0], F.x, R[1] * F.y);
    ty += fmaf(R[4], F.x, R[5] * F.y);
  }
  const float fk = (float)k;
  avg[i] = make_float4(tx / fk, ty / fk, 0.f, 0.f);
}

hipError_t launch_mog2d_deposit(const float4* pos, long long n, int k, int lg, float W,
                                const float* rot, const float* off, const float* frot, float lim,
                                long long* grids, hipStream_t s) {
  const int g = 1 << lg, cells = g * g;
  const float h = W / (float)g;
  hipError_t e = launch_fill_zero(grids, (long long)k * cells, s);
  if (e != hipSuccess || n <= 0) return e;
  const long long chunk = 16384;
  const long long chunks = (n + chunk - 1) / chunk;
  hipLaunchKernelGGL(deposit2d_kernel, dim3((unsigned)chunks, (unsigned)k), dim3(256),
                     (size_t)cells * sizeof(unsigned int), s, pos, n, chunk, rot, off, frot, W, h,
                     lg, lim, (unsigned long long*)grids);
  return hipGetLastError();
}

hipError_t launch_mog2d_forces(const float4* pos, long long n, int k, int lg, float W,
                               const float* rot, const float* off, const float* frot, float lim,
                               const float2* tw, const float* gtab, const float* kd,
                               long long* grids, float2* force, float4* avg, hipStream_t s) {
  const int g = 1 << lg, cells = g * g;
  const float h = W / (float)g;
  hipError_t e = launch_mog2d_deposit(pos, n, k, lg, W, rot, off, frot, lim, grids, s);
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)(g + cells / 2 + 1 + 2 * cells) * sizeof(float2);
  if ((e = hipFuncSetAttribute((const void*)solve2d_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(solve2d_kernel, dim3((unsigned)k), dim3(256), lds, s, grids, tw, gtab, kd, lg,
                     force);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (n > 0) {
    hipLaunchKernelGGL(gather2d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos, n,
                       k, rot, off, frot, W, h, lg, lim, force, avg);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace mog
