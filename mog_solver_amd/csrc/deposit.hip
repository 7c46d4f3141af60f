// NGP mass deposit of the MOG step (replaces shaders/mass.glsl:27-55 and the
// host-side clear of main.py:558-562).
//
// Work is tiled as (particle chunk) x (group of G grids).  A workgroup keeps
// a private 32-bit histogram of each of its G grids in LDS, streams its
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

// Deferred (particle, grid) pairs per workgroup: pairs that the fast cell
// test cannot decide (near a face, or a particle outside the fast bound).
constexpr int kDeferCap = 512;

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ unsigned int lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned int)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned int)mask, 0u));
}

// LDS-privatised deposit; grid (chunks, ceil(k / G)).  Dynamic LDS holds
// G * cells uint32 counters.  The G fast rows are wave-uniform and live in
// scalar registers for the whole chunk.
//
// Every lane adds `decided ? milli : 0` to its fast cell (always in range);
// the pairs the fast test cannot decide (about 1e-3 of them) are appended to
// an LDS queue, and after the particle loop whole waves run the exact
// float32 sequence for the queued pairs.  The loop itself then makes no
// function call, so the next iteration's particle loads stay in flight
// across it (a call would wait for them), and no wave runs the exact
// sequence for a single lane.  A full queue falls back to deciding in place.
// Particles whose milli-mass does not fit the 32-bit LDS partials take the
// 64-bit global path (rare: huge masses).
template <int G>
__global__ __launch_bounds__(1024) void deposit_lds_kernel(
    const float4* __restrict__ pos, long long n, long long chunk,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    Params P, unsigned long long lds_limit, unsigned long long* __restrict__ grids) {
  extern __shared__ unsigned int hist[];
  __shared__ unsigned int defer_p[kDeferCap];  // particle offset within the chunk
  __shared__ unsigned char defer_g[kDeferCap];  // grid within the group
  __shared__ unsigned int defer_n;
  const int cells = P.cells;
  // tiles in chunk-major order: the groups of one chunk share an XCD's L2
  const int groups = (P.k + G - 1) / G;
  const unsigned int tile = xcd_tile(blockIdx.x, gridDim.x);
  const long long chunk_id = tile / groups;
  const int grid0 = (int)(tile % groups) * G;
  const int ng = min(G, P.k - grid0);
  const int total = ng * cells;
  for (int i = threadIdx.x; i < total; i += blockDim.x) hist[i] = 0u;
  if (threadIdx.x == 0) defer_n = 0u;

  float F[G][12];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int gr = min(grid0 + gi, P.k - 1);
#pragma unroll
    for (int j = 0; j < 12; ++j) F[gi][j] = frot[12 * gr + j];
#pragma unroll
    for (int m = 0; m < 3; ++m) F[gi][4 * m + 3] = to_vgpr(F[gi][4 * m + 3]);
  }
  __syncthreads();

  const long long p0 = chunk_id * chunk;
  const long long p1 = min(n, p0 + chunk);
  const unsigned int lds_limit32 =
      (unsigned int)(lds_limit < 0xFFFFFFFFull ? lds_limit : 0xFFFFFFFFull);
  // UP particles per thread per iteration: their loads are issued together.
  constexpr int UP = 2;
  const long long step = (long long)UP * blockDim.x;
  // the next iteration's particles are loaded while this one computes
  float4 qn[UP];
#pragma unroll
  for (int u = 0; u < UP; ++u) {
    const long long p = p0 + threadIdx.x + (long long)u * blockDim.x;
    qn[u] = p < p1 ? pos[p] : make_float4(0.f, 0.f, 0.f, 0.f);  // mass 0 is skipped
  }
  for (long long pb = p0 + threadIdx.x; pb < p1; pb += step) {
    float4 q[UP];
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      q[u] = qn[u];
      const long long p = pb + step + (long long)u * blockDim.x;
      qn[u] = p < p1 ? pos[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    // all UP x G cells first, with no branch between them (independent
    // work the scheduler can interleave), then the atomics, then one
    // wave-uniform test for the rare paths
    unsigned int m32[UP];
    bool small[UP], big[UP];
    float cc[UP][3];
    int pf[UP];
    int cell[UP][G];
    bool ok[UP][G];
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      // uint(m * 1000) (milli_of): 0 exactly when the float product is below
      // 1 (or negative, or NaN); while it is below 2^32 a 32-bit truncation
      // gives it, which spares the 64-bit conversion for every real mass
      const float mm = q[u].w * 1000.0f;
      const bool live = mm >= 1.0f;
      m32[u] = live && mm < 4294967296.0f ? (unsigned int)mm : 0u;
      small[u] = live && mm < 4294967296.0f && m32[u] <= lds_limit32;
      big[u] = live && !small[u];
      cc[u][0] = q[u].x - P.half;
      cc[u][1] = q[u].y - P.half;
      cc[u][2] = q[u].z - P.half;
      pf[u] = particle_fast(cc[u][0], cc[u][1], cc[u][2], P.CB);
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
        ok[u][gi] = cell_fast_a(cc[u][0], cc[u][1], cc[u][2], F[gi], F[gi][3], F[gi][7],
                                F[gi][11], P.lim, P.lg, &cell[u][gi]) & pf[u];
    }
    // grids gi >= ng (last group) add 0 to allocated, unflushed counters
    bool rare = false;
#pragma unroll
    for (int u = 0; u < UP; ++u) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
        const bool take = gi < ng && small[u];
        // the addend is hidden from the optimiser: a known-zero add would be
        // turned into an atomic load and a wait
        unsigned int v = take && ok[u][gi] ? m32[u] : 0u;
        asm volatile("" : "+v"(v));
        atomicAdd(&hist[gi * cells + cell[u][gi]], v);
        rare |= take && !ok[u][gi];
      }
      rare |= big[u];
    }
    if (__ballot(rare)) {  // wave-uniform, rare
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const float c0 = cc[u][0], c1 = cc[u][1], c2 = cc[u][2];
        const long long p = pb + (long long)u * blockDim.x;
        if (big[u]) {  // huge mass: the 64-bit global grids
          const unsigned long long milli = milli_of(q[u].w);
#pragma unroll
          for (int gi = 0; gi < G; ++gi) {
            if (gi < ng) {
              const int gr = grid0 + gi;
              const int c = cell_index(c0, c1, c2, pf[u], F[gi], rot + 12 * gr, off + 4 * gr, P);
              atomicAdd(&grids[(long long)gr * cells + c], milli);
            }
          }
        }
#pragma unroll
        for (int gi = 0; gi < G; ++gi) {
          const bool defer = gi < ng && small[u] && !ok[u][gi];
          const unsigned long long need = __ballot(defer);
          if (need) {  // queue the undecided pairs: one reservation per wave
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((long long)need) - 1;
            unsigned int base = 0u;
            if (lane == leader) base = atomicAdd(&defer_n, (unsigned int)__popcll(need));
            base = __builtin_amdgcn_readlane(base, leader);
            if (defer) {
              const unsigned int slot = base + lanes_below(need);
              if (slot < (unsigned int)kDeferCap) {
                defer_p[slot] = (unsigned int)(p - p0);
                defer_g[slot] = (unsigned char)gi;
              } else {
                const int gr = grid0 + gi;
                const int ex = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W,
                                               P.half, P.h, P.lg);
                atomicAdd(&hist[gi * cells + ex], m32[u]);
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  // the deferred pairs, one per lane
  const unsigned int nd = min(defer_n, (unsigned int)kDeferCap);
  for (unsigned int d = threadIdx.x; d < nd; d += blockDim.x) {
    const int gi = defer_g[d];
    const int gr = grid0 + gi;
    const float4 q = pos[p0 + defer_p[d]];
    const int cell = cell_exact_call(q.x - P.half, q.y - P.half, q.z - P.half, rot + 12 * gr,
                                     off + 4 * gr, P.W, P.half, P.h, P.lg);
    atomicAdd(&hist[gi * cells + cell], (unsigned int)(q.w * 1000.0f));
  }
  __syncthreads();
  unsigned long long* dst = grids + (long long)grid0 * cells;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const unsigned int v = hist[i];
    if (v) atomicAdd(&dst[i], (unsigned long long)v);
  }
}

// Direct global deposit for grids too large to privatise in LDS (g = 64).
__global__ __launch_bounds__(256) void deposit_global_kernel(
    const float4* __restrict__ pos, long long n, int G, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot, Params P,
    unsigned long long* __restrict__ grids) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int grid0 = blockIdx.y * G;
  const int ng = min(G, P.k - grid0);
  const float4 q = pos[p];
  const unsigned long long milli = milli_of(q.w);
  if (milli == 0ull) return;
  const float c0 = q.x - P.half;
  const float c1 = q.y - P.half;
  const float c2 = q.z - P.half;
  const int pfast = particle_fast(c0, c1, c2, P.CB);
  for (int gi = 0; gi < ng; ++gi) {
    const int gr = grid0 + gi;
    const int cell = cell_index(c0, c1, c2, pfast, frot + 12 * gr, rot + 12 * gr, off + 4 * gr, P);
    atomicAdd(&grids[(long long)gr * P.cells + cell], milli);
  }
}

// Sum of v over the 64 lanes of the wave (every lane must be active):
// DPP within rows of 16 (quad swaps, half-row and row mirrors), then the four
// row sums through readlane.
__device__ __forceinline__ unsigned wave_sum(unsigned v) {
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
  return (unsigned)__builtin_amdgcn_readlane((int)v, 0) + (unsigned)__builtin_amdgcn_readlane((int)v, 16) +
         (unsigned)__builtin_amdgcn_readlane((int)v, 32) + (unsigned)__builtin_amdgcn_readlane((int)v, 48);
}

// hist[cell] += v over the wave (v == 0: no contribution).  Lanes that share
// the cell of a leader lane are summed and added with one atomic (two
// rounds); the rest add their own value.  Called with every lane active.
__device__ __forceinline__ void wave_aggregate_add(unsigned int* hist, int cell, unsigned v) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(v != 0u);
#pragma unroll
  for (int round = 0; round < 2; ++round) {
    if (pending == 0ull) return;
    const int leader = __ffsll((long long)pending) - 1;
    const int lc = __builtin_amdgcn_readlane(cell, leader);
    const bool mine = ((pending >> lane) & 1ull) && cell == lc;
    const unsigned s = wave_sum(mine ? v : 0u);
    if (lane == leader) atomicAdd(&hist[lc], s);
    pending &= ~__ballot(mine);
  }
  if ((pending >> lane) & 1ull) atomicAdd(&hist[cell], v);
}

// Deposit over positions in Morton order (spos[t] = pos[order[t]]): the
// lanes of a wave are neighbours in space and mostly share cells, so LDS
// adds are aggregated per wave.  With CACHE the cell of every (grid, slot)
// pair is stored in cells_out[cache_index(t, grid, k)] for the gather.
template <int G, bool CACHE>
__global__ __launch_bounds__(1024) void deposit_sorted_kernel(
    const float4* __restrict__ spos, long long n, long long chunk,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    Params P, unsigned long long lds_limit, unsigned long long* __restrict__ grids,
    unsigned short* __restrict__ cells_out) {
  extern __shared__ unsigned int hist[];
  const int cells = P.cells;
  const int groups = (P.k + G - 1) / G;
  const unsigned int tile = xcd_tile(blockIdx.x, gridDim.x);
  const long long chunk_id = tile / groups;
  const int grid0 = (int)(tile % groups) * G;
  const int ng = min(G, P.k - grid0);
  const int total = ng * cells;
  for (int i = threadIdx.x; i < total; i += blockDim.x) hist[i] = 0u;
  float F[G][12];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int gr = min(grid0 + gi, P.k - 1);
#pragma unroll
    for (int j = 0; j < 12; ++j) F[gi][j] = frot[12 * gr + j];
  }
  __syncthreads();
  const long long p0 = chunk_id * chunk;
  const long long p1 = min(n, p0 + chunk);
  for (long long base = p0; base < p1; base += blockDim.x) {  // same trip count for all lanes
    const long long p = base + threadIdx.x;
    const bool live = p < p1;
    const float4 q = live ? spos[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const unsigned long long milli = milli_of(q.w);
    const unsigned lds_v = (milli <= lds_limit) ? (unsigned)milli : 0u;
    const bool glob = milli > lds_limit;
    const float c0 = q.x - P.half;
    const float c1 = q.y - P.half;
    const float c2 = q.z - P.half;
    const int pfast = particle_fast(c0, c1, c2, P.CB);
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      if (gi < ng) {
        const int gr = grid0 + gi;
        const int cell = cell_index(c0, c1, c2, pfast, F[gi], rot + 12 * gr, off + 4 * gr, P);
        if (CACHE && live) cells_out[cache_index(p, gr, P.k)] = (unsigned short)cell;
        if (glob) atomicAdd(&grids[(long long)gr * cells + cell], milli);
        wave_aggregate_add(hist + gi * cells, cell, lds_v);
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

__global__ __launch_bounds__(256) void permute_kernel(const float4* __restrict__ pos,
                                                      const int* __restrict__ order, long long n,
                                                      float4* __restrict__ spos) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) spos[t] = pos[order[t]];
}

hipError_t launch_permute(const float4* pos, const int* order, long long n, float4* spos,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(permute_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                     order, n, spos);
  return hipGetLastError();
}

static const size_t kMaxLds = 160 * 1024;

template <int G>
static hipError_t launch_lds(const Params& P, const float4* pos, long long n, long long chunks,
                             long long chunk, int threads, const float* rot, const float* off,
                             const float* frot, unsigned long long* grids,
                             unsigned short* cells_out, bool sorted, hipStream_t s) {
  const int groups = (P.k + G - 1) / G;
  const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)chunk;
  const size_t lds = (size_t)G * P.cells * 4;
  const void* fn = !sorted ? (const void*)deposit_lds_kernel<G>
                   : cells_out ? (const void*)deposit_sorted_kernel<G, true>
                               : (const void*)deposit_sorted_kernel<G, false>;
  // the exact dynamic size: the kernels also hold static LDS
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)(chunks * groups));
  if (!sorted)
    hipLaunchKernelGGL(deposit_lds_kernel<G>, grid, dim3(threads), lds, s, pos, n, chunk, rot, off,
                       frot, P, lds_limit, grids);
  else if (cells_out)
    hipLaunchKernelGGL((deposit_sorted_kernel<G, true>), grid, dim3(threads), lds, s, pos, n, chunk,
                       rot, off, frot, P, lds_limit, grids, cells_out);
  else
    hipLaunchKernelGGL((deposit_sorted_kernel<G, false>), grid, dim3(threads), lds, s, pos, n,
                       chunk, rot, off, frot, P, lds_limit, grids, cells_out);
  return hipGetLastError();
}

hipError_t launch_deposit(const Params& P, const float4* pos, long long n, const float* rot,
                          const float* off, const float* frot, unsigned long long* grids,
                          const DepositTune& t, hipStream_t s) {
  return launch_deposit_ex(P, pos, n, rot, off, frot, grids, nullptr, false, t, s);
}

hipError_t launch_deposit_ex(const Params& P, const float4* pos, long long n, const float* rot,
                             const float* off, const float* frot, unsigned long long* grids,
                             unsigned short* cells_out, bool sorted, const DepositTune& t,
                             hipStream_t s) {
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
  // G in {1, 2, 4, 8}, within the LDS and at most k
  int G = 1;
  while (G * 2 <= t.grids_per_wg && G * 2 <= 8 && (size_t)(G * 2) * grid_bytes <= kMaxLds &&
         G * 2 <= P.k)
    G *= 2;
  const int groups = (P.k + G - 1) / G;
  const int threads = t.threads;
  long long chunks = (t.wg_target + groups - 1) / groups;
  // one-dimensional launch of chunks x groups workgroups: under 2^32 lanes
  const long long max_chunks = 0xFFFFFFFFll / ((long long)threads * groups);
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  long long chunk = (n + chunks - 1) / chunks;
  if (chunk < threads) chunk = threads;
  chunks = (n + chunk - 1) / chunk;
  switch (G) {
    case 8:
      return launch_lds<8>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, s);
    case 4:
      return launch_lds<4>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, s);
    case 2:
      return launch_lds<2>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, s);
    default:
      return launch_lds<1>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, s);
  }
}

}  // namespace mog
