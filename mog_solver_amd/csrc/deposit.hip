// NGP mass deposit of the MOG step (replaces shaders/mass.glsl:27-55 and the
// host-side clear of main.py:558-562).
//
// Work is tiled as (particle chunk) x (group of grids).  A workgroup keeps a
// private 32-bit histogram of each grid of its group in LDS, streams its
// particle chunk once, and for every (particle, grid) pair computes the
// reference's cell (cellmath.h: three FMAs and a face-distance test, with the
// exact float32 sequence as fallback near faces) and adds the particle's
// milli-mass with an LDS atomic.  At the end the non-zero LDS cells are added
// to the int64 density grids in HBM with 64-bit atomics.  Integer addition is
// associative, so the grids are bit-exact and independent of the order in
// which atomics land.
//
// 32-bit LDS partials are exact while a chunk's sum stays below 2^32: a
// particle whose milli-mass exceeds (2^32-1)/chunk goes straight to the
// 64-bit global grid instead, so no LDS cell can overflow.
#include "mog_common.h"

namespace mog {

__global__ void fill_zero_kernel(long long* __restrict__ p, long long count) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = 0;
}

hipError_t launch_fill_zero(long long* p, long long count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  long long blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, count);
  return hipGetLastError();
}

// LDS-privatised deposit.  grid (chunks, grid_groups); dynamic LDS holds
// ng * cells uint32 counters.  With CACHE the cell of every (grid, particle)
// pair is also stored as uint16 in cells_out[grid * n + p] for the gather of
// the same step (same positions, so the same cells).
template <bool CACHE>
__global__ __launch_bounds__(1024) void deposit_lds_kernel(
    const float4* __restrict__ pos, long long n, long long chunk, int G,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    Params P, unsigned long long lds_limit, unsigned long long* __restrict__ grids,
    unsigned short* __restrict__ cells_out) {
  extern __shared__ unsigned int hist[];
  const int cells = P.cells;
  const int grid0 = blockIdx.y * G;
  const int ng = min(G, P.k - grid0);
  const int total = ng * cells;
  for (int i = threadIdx.x; i < total; i += blockDim.x) hist[i] = 0u;
  __syncthreads();

  const long long p0 = (long long)blockIdx.x * chunk;
  const long long p1 = min(n, p0 + chunk);
  for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const float4 q = pos[p];
    const float c0 = q.x - P.half;
    const float c1 = q.y - P.half;
    const float c2 = q.z - P.half;
    const unsigned long long milli = milli_of(q.w);
    if (milli == 0ull && !CACHE) continue;
    const bool pfast = particle_fast(c0, c1, c2, P.CB);
    if (milli <= lds_limit) {
      const unsigned int mv = (unsigned int)milli;
      for (int gi = 0; gi < ng; ++gi) {
        const int grid = grid0 + gi;
        const int cell =
            cell_index(c0, c1, c2, pfast, frot + 12 * grid, rot + 12 * grid, off + 4 * grid, P);
        atomicAdd(&hist[gi * cells + cell], mv);
        if (CACHE) cells_out[(long long)grid * n + p] = (unsigned short)cell;
      }
    } else {
      for (int gi = 0; gi < ng; ++gi) {
        const int grid = grid0 + gi;
        const int cell =
            cell_index(c0, c1, c2, pfast, frot + 12 * grid, rot + 12 * grid, off + 4 * grid, P);
        atomicAdd(&grids[(long long)grid * cells + cell], milli);
        if (CACHE) cells_out[(long long)grid * n + p] = (unsigned short)cell;
      }
    }
  }
  __syncthreads();
  unsigned long long* dst = grids + (long long)grid0 * cells;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const unsigned int v = hist[i];
    if (v) atomicAdd(&dst[i], (unsigned long long)v);
  }
}

// Direct global deposit for grids too large to privatise in LDS.
__global__ __launch_bounds__(256) void deposit_global_kernel(
    const float4* __restrict__ pos, long long n, int G, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot, Params P,
    unsigned long long* __restrict__ grids) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int grid0 = blockIdx.y * G;
  const int ng = min(G, P.k - grid0);
  const float4 q = pos[p];
  const float c0 = q.x - P.half;
  const float c1 = q.y - P.half;
  const float c2 = q.z - P.half;
  const unsigned long long milli = milli_of(q.w);
  if (milli == 0ull) return;
  const bool pfast = particle_fast(c0, c1, c2, P.CB);
  for (int gi = 0; gi < ng; ++gi) {
    const int grid = grid0 + gi;
    const int cell =
        cell_index(c0, c1, c2, pfast, frot + 12 * grid, rot + 12 * grid, off + 4 * grid, P);
    atomicAdd(&grids[(long long)grid * P.cells + cell], milli);
  }
}

static const size_t kMaxLds = 160 * 1024;

hipError_t launch_deposit(const Params& P, const float4* pos, long long n, const float* rot,
                          const float* off, const float* frot, unsigned long long* grids,
                          unsigned short* cells_out, const DepositTune& t, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t grid_bytes = (size_t)P.cells * 4;
  if (grid_bytes > kMaxLds) {
    const int G = 8;
    const int groups = (P.k + G - 1) / G;
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(deposit_global_kernel, dim3((unsigned)blocks, (unsigned)groups), dim3(256),
                       0, s, pos, n, G, rot, off, frot, P, grids);
    return hipGetLastError();
  }
  int G = t.grids_per_wg;
  const int gmax = (int)(kMaxLds / grid_bytes);
  if (G > gmax) G = gmax;
  if (G > P.k) G = P.k;
  if (G < 1) G = 1;
  const int groups = (P.k + G - 1) / G;
  int threads = t.threads;
  long long chunks = (t.wg_target + groups - 1) / groups;
  if (chunks < 1) chunks = 1;
  long long chunk = (n + chunks - 1) / chunks;
  if (chunk < threads) chunk = threads;
  chunks = (n + chunk - 1) / chunk;
  const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)chunk;
  const size_t lds = (size_t)G * grid_bytes;
  const void* fn = cells_out ? (const void*)deposit_lds_kernel<true> : (const void*)deposit_lds_kernel<false>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  if (e != hipSuccess) return e;
  if (cells_out)
    hipLaunchKernelGGL(deposit_lds_kernel<true>, dim3((unsigned)chunks, (unsigned)groups),
                       dim3(threads), lds, s, pos, n, chunk, G, rot, off, frot, P, lds_limit, grids,
                       cells_out);
  else
    hipLaunchKernelGGL(deposit_lds_kernel<false>, dim3((unsigned)chunks, (unsigned)groups),
                       dim3(threads), lds, s, pos, n, chunk, G, rot, off, frot, P, lds_limit, grids,
                       cells_out);
  return hipGetLastError();
}

}  // namespace mog
