// FP64 diagnostics of a particle state: momentum, kinetic energy, mass and
// the grid potential energy of the MOG ensemble.  The reference has no
// diagnostics (it prints its configuration only, main.py:147-155); the
// potential is the one whose gradient the step's force grids hold
// (greens.glsl:40-47: phi_hat = -2 pi G rho_hat / k^2 where the float32
// k2 > 1e-4, DC zeroed; rho = milli-mass / 1000 per cell, complex.glsl:23).
//
// Potential energy of one grid, by Parseval for the unnormalised DFT:
//   U = 1/2 sum_c rho(c) phi(c) = sum_k T(k) |m_hat(k)|^2,
//   T(k) = G_hat(k) / (2 g^3 10^6),  m_hat = DFT of the int64 milli-mass.
// Every operation is FP64.  m is real, so m_hat(-k) = conj m_hat(k) and
// only kx in [0, g/2] is transformed; kx = 0 and g/2 count once, the others
// twice (T depends on k^2 only).  One workgroup per (grid, kx): the x
// transform of the grid for that kx (a g x g plane), then y and z in LDS,
// then the weighted sum.  All sums are fixed-order tree reductions, so the
// result does not depend on scheduling.
#include "mog_common.h"

namespace mog {

constexpr int kDiagThreads = 256;
constexpr int kDiagBlocks = 1024;  // particle-sum partials

// Fixed-order sum of the workgroup's values (kDiagThreads lanes).
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// partial[b * 5 + j], j = (sum m vx, sum m vy, sum m vz, sum m |v|^2, sum m)
__global__ __launch_bounds__(kDiagThreads) void particle_sums_kernel(
    const float4* __restrict__ pos, const float4* __restrict__ vel, long long n,
    double* __restrict__ partial) {
  __shared__ double red[kDiagThreads];
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kDiagThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kDiagThreads) {
    const double m = (double)pos[i].w;
    const float4 v = vel[i];
    const double vx = v.x, vy = v.y, vz = v.z;
    a[0] += m * vx;
    a[1] += m * vy;
    a[2] += m * vz;
    a[3] += m * (vx * vx + vy * vy + vz * vz);
    a[4] += m;
  }
  for (int j = 0; j < 5; ++j) {
    const double s = block_sum(a[j], red);
    if (threadIdx.x == 0) partial[blockIdx.x * 5 + j] = s;
  }
}

// One workgroup per (grid, kx).  tw[j] = exp(-2 pi i j / g) (FP64), tab the
// (g,g,g) [kz][ky][kx] table T(k); egrid[grid * (g/2+1) + kx] = the
// weighted sum of T |m_hat|^2 over (ky, kz) for this kx.
__global__ __launch_bounds__(kDiagThreads) void grid_energy_kernel(
    const long long* __restrict__ grids, int g, const double2* __restrict__ tw,
    const double* __restrict__ tab, double* __restrict__ egrid) {
  extern __shared__ double2 plane[];  // 2 g^2: A[z][y], then B[z][ky]
  __shared__ double red[kDiagThreads];
  const int nkx = g / 2 + 1;
  const int grid = blockIdx.x / nkx;
  const int kx = blockIdx.x % nkx;
  const int gm = g - 1;
  const int g2 = g * g;
  double2* A = plane;
  double2* B = plane + g2;
  const long long* m = grids + (long long)grid * g2 * g;
  // A[z][y] = sum_x m[z][y][x] w^(kx x)
  for (int zy = threadIdx.x; zy < g2; zy += kDiagThreads) {
    const long long* row = m + (long long)zy * g;
    double re = 0.0, im = 0.0;
    for (int x = 0; x < g; ++x) {
      const double v = (double)row[x];
      const double2 w = tw[(kx * x) & gm];
      re += v * w.x;
      im += v * w.y;
    }
    A[zy] = make_double2(re, im);
  }
  __syncthreads();
  // B[z][ky] = sum_y A[z][y] w^(ky y)
  for (int zk = threadIdx.x; zk < g2; zk += kDiagThreads) {
    const int z = zk / g, ky = zk % g;
    double re = 0.0, im = 0.0;
    for (int y = 0; y < g; ++y) {
      const double2 a = A[z * g + y];
      const double2 w = tw[(ky * y) & gm];
      re += a.x * w.x - a.y * w.y;
      im += a.x * w.y + a.y * w.x;
    }
    B[zk] = make_double2(re, im);
  }
  __syncthreads();
  // sum over (ky, kz) of T |sum_z B[z][ky] w^(kz z)|^2
  double acc = 0.0;
  for (int kk = threadIdx.x; kk < g2; kk += kDiagThreads) {
    const int kz = kk / g, ky = kk % g;
    double re = 0.0, im = 0.0;
    for (int z = 0; z < g; ++z) {
      const double2 b = B[z * g + ky];
      const double2 w = tw[(kz * z) & gm];
      re += b.x * w.x - b.y * w.y;
      im += b.x * w.y + b.y * w.x;
    }
    acc += tab[((long long)kz * g + ky) * g + kx] * (re * re + im * im);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) egrid[blockIdx.x] = (kx == 0 || 2 * kx == g) ? s : 2.0 * s;
}

// out[0..2] momentum, out[3] kinetic, out[4] potential (mean over grids),
// out[5] mass, out[6..7] = 0; sums in fixed order.
__global__ __launch_bounds__(kDiagThreads) void diag_finalize_kernel(
    const double* __restrict__ partial, int nblocks, const double* __restrict__ egrid, int k,
    int nkx, double* __restrict__ out) {
  __shared__ double red[kDiagThreads];
  const int t = threadIdx.x;
  if (t < 5) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[b * 5 + t];
    red[t] = s;
  }
  // per grid: its kx terms in order (one lane per grid), then the mean
  double u = 0.0;
  for (int gr = t; gr < k; gr += kDiagThreads) {
    double e = 0.0;
    for (int kx = 0; kx < nkx; ++kx) e += egrid[(long long)gr * nkx + kx];
    u += e;
  }
  __syncthreads();
  double sums[5];
  for (int j = 0; j < 5; ++j) sums[j] = red[j];
  __syncthreads();
  const double U = block_sum(u, red);
  if (t == 0) {
    out[0] = sums[0];
    out[1] = sums[1];
    out[2] = sums[2];
    out[3] = 0.5 * sums[3];
    out[4] = U / (double)k;
    out[5] = sums[4];
    out[6] = 0.0;
    out[7] = 0.0;
  }
}

size_t diag_scratch_doubles(int k, int g) {
  return (size_t)kDiagBlocks * 5 + (size_t)k * (g / 2 + 1);
}

hipError_t launch_diagnostics(const float4* pos, const float4* vel, long long n,
                              const long long* grids, int k, int g, const double2* tw,
                              const double* tab, double* scratch, double* out, hipStream_t s) {
  double* partial = scratch;
  double* egrid = scratch + (size_t)kDiagBlocks * 5;
  const int nkx = g / 2 + 1;
  long long blocks = (n + kDiagThreads - 1) / kDiagThreads;
  if (blocks > kDiagBlocks) blocks = kDiagBlocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(particle_sums_kernel, dim3((unsigned)blocks), dim3(kDiagThreads), 0, s, pos,
                     vel, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = 2 * (size_t)g * g * sizeof(double2);
  e = hipFuncSetAttribute((const void*)grid_energy_kernel,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(grid_energy_kernel, dim3((unsigned)(k * nkx)), dim3(kDiagThreads), lds, s,
                     grids, g, tw, tab, egrid);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(diag_finalize_kernel, dim3(1), dim3(kDiagThreads), 0, s, partial,
                     (int)blocks, egrid, k, nkx, out);
  return hipGetLastError();
}

}  // namespace mog
