// Batched ensemble of independent MOG simulations: deposit and gather+update
// as one persistent launch each over all simulations (the solve's persistent
// passes live beside the single-simulation passes in spectral.hip).
//
// Deposit: the host cuts every simulation into tasks (particle chunk x group
// of grids whose histograms fit the LDS budget) and orders them by cost,
// largest first.  A workgroup takes the next task from a device counter,
// clears its LDS histograms, streams the chunk, and flushes non-zero cells
// with 64-bit atomics, exactly as the single-simulation deposit does, so the
// grids are the same integers.  Dynamic task pickup balances the uneven
// simulations (n spans 50x, k 7.5x, g^3 8x).
//
// Gather+update: a grid-stride loop over blocks of particles that never
// straddle two simulations; a block finds its simulation by binary search
// over the block offsets and then runs the single-simulation world-frame
// gather (same summation order, same integrator), so the new state is
// bit-identical to stepping each simulation alone.
#include "mog_common.h"

namespace mog {

// Deferred (particle, grid) pairs per task: pairs that the fast cell test
// cannot decide, settled by whole waves at the end of the task.
constexpr int kBatchDeferCap = 512;

// One particle's inputs to a deposit group.
struct DepositParticle {
  float c0, c1, c2;    // centred position
  int pfast;           // fast-path admission
  bool small, big;     // 32-bit LDS partial / 64-bit global add
  unsigned int m32;    // milli-mass (small)
  float mass;
  unsigned int pofs;   // offset within the task
};

// A group of UP particles x GU grids (gi0 .. gi0+GU-1, all < ng): the UP x GU
// fast cells with no branch between them, their atomics, then one
// wave-uniform test for the rare paths (undecided pairs go to the task's
// queue; huge masses to the 64-bit global grids).
template <int UP, int GU>
__device__ __forceinline__ void deposit_group(
    int gi0, int gbase, int cells, const Params& P, const DepositParticle (&q)[UP],
    const float* __restrict__ frot, const float* __restrict__ rot, const float* __restrict__ off,
    unsigned int* hist, unsigned long long* __restrict__ dst, unsigned int* defer_p,
    unsigned short* defer_g, unsigned int* defer_n) {
  int cell[UP][GU];
  bool ok[UP][GU];
  bool rare = false;
#pragma unroll
  for (int a = 0; a < UP; ++a) {
    rare |= q[a].big;
#pragma unroll
    for (int u = 0; u < GU; ++u)
      ok[a][u] = cell_fast(q[a].c0, q[a].c1, q[a].c2, frot + 12 * (gbase + gi0 + u), P.lim, P.lg,
                           &cell[a][u]) &
                 q[a].pfast;
  }
#pragma unroll
  for (int a = 0; a < UP; ++a)
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      // hidden from the optimiser: a known-zero add would become an atomic
      // load and a wait
      unsigned int v = q[a].small && ok[a][u] ? q[a].m32 : 0u;
      asm volatile("" : "+v"(v));
      atomicAdd(&hist[(gi0 + u) * cells + cell[a][u]], v);
      rare |= q[a].small && !ok[a][u];
    }
  if (!__ballot(rare)) return;  // wave-uniform
#pragma unroll
  for (int a = 0; a < UP; ++a)
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      const int gi = gi0 + u;
      const int gr = gbase + gi;
      if (q[a].big) {
        const int c = ok[a][u] ? cell[a][u]
                               : cell_exact_call(q[a].c0, q[a].c1, q[a].c2, rot + 12 * gr,
                                                 off + 4 * gr, P.W, P.half, P.h, P.lg);
        atomicAdd(&dst[(long long)gi * cells + c], milli_of(q[a].mass));
      }
      const bool defer = q[a].small && !ok[a][u];
      const unsigned long long need = __ballot(defer);
      if (need) {  // queue: one reservation per wave
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)need) - 1;
        unsigned int base = 0u;
        if (lane == leader) base = atomicAdd(defer_n, (unsigned int)__popcll(need));
        base = __builtin_amdgcn_readlane(base, leader);
        if (defer) {
          const unsigned int slot =
              base + __builtin_amdgcn_mbcnt_hi((unsigned int)(need >> 32),
                                               __builtin_amdgcn_mbcnt_lo((unsigned int)need, 0u));
          if (slot < (unsigned int)kBatchDeferCap) {
            defer_p[slot] = q[a].pofs;
            defer_g[slot] = (unsigned short)gi;
          } else {
            const int ex = cell_exact_call(q[a].c0, q[a].c1, q[a].c2, rot + 12 * gr, off + 4 * gr,
                                           P.W, P.half, P.h, P.lg);
            atomicAdd(&hist[gi * cells + ex], q[a].m32);
          }
        }
      }
    }
}

// Inputs of the particle at `p` (mass 0 when dead).
__device__ __forceinline__ DepositParticle deposit_particle(const float4 qv, bool in, long long p,
                                                            long long p0, const Params& P,
                                                            unsigned int lds_limit32) {
  // milli-mass as in the single-simulation deposit: 32-bit LDS partials
  // while they cannot overflow, the 64-bit global grids otherwise
  DepositParticle d;
  const float mm = in ? qv.w * 1000.0f : 0.0f;
  const bool live = mm >= 1.0f;
  d.m32 = live && mm < 4294967296.0f ? (unsigned int)mm : 0u;
  d.small = live && mm < 4294967296.0f && d.m32 <= lds_limit32;
  d.big = live && !d.small;
  d.mass = qv.w;
  d.c0 = qv.x - P.half;
  d.c1 = qv.y - P.half;
  d.c2 = qv.z - P.half;
  d.pfast = particle_fast(d.c0, d.c1, d.c2, P.CB);
  d.pofs = (unsigned int)(p - p0);
  return d;
}

__global__ __launch_bounds__(1024) void batch_deposit_kernel(
    const float4* __restrict__ pos, const BatchSim* __restrict__ sims,
    const DepositTask* __restrict__ tasks, int ntasks, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    unsigned long long* __restrict__ grids, unsigned long long* __restrict__ counter) {
  extern __shared__ unsigned int hist[];
  __shared__ int cur;
  __shared__ unsigned int defer_p[kBatchDeferCap];  // particle offset within the task
  __shared__ unsigned short defer_g[kBatchDeferCap];  // grid within the task
  __shared__ unsigned int defer_n;
  for (;;) {
    if (threadIdx.x == 0) {
      cur = (int)atomicAdd(counter, 1ull);
      defer_n = 0u;
    }
    __syncthreads();
    const int ti = cur;
    if (ti >= ntasks) break;  // uniform across the workgroup
    const DepositTask task = tasks[ti];
    const BatchSim& b = sims[task.sim];
    const Params P = b.P;
    const int cells = P.cells;
    const int ng = task.ng;
    const int total = ng * cells;
    for (int i = threadIdx.x; i < total; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const long long p0 = b.p_off + task.p0, p1 = b.p_off + task.p1;
    const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)(task.p1 - task.p0);
    const unsigned int lds_limit32 =
        (unsigned int)(lds_limit < 0xFFFFFFFFull ? lds_limit : 0xFFFFFFFFull);
    const int gbase = b.grid_off + task.grid0;
    unsigned long long* dst = grids + b.cell_off + (long long)task.grid0 * cells;
    if (ng == 1) {
      // one grid: two particles per lane and iteration, the next two loaded
      // while these are deposited
      const long long stride = 2LL * blockDim.x;
      long long p = p0 + threadIdx.x;
      float4 n0 = p < p1 ? pos[p] : make_float4(0.f, 0.f, 0.f, 0.f);
      float4 n1 = p + blockDim.x < p1 ? pos[p + blockDim.x] : make_float4(0.f, 0.f, 0.f, 0.f);
      for (; p < p1; p += stride) {  // lane-varying trip count: no barrier inside
        const float4 a = n0, b2 = n1;
        const long long pb = p + blockDim.x;
        n0 = p + stride < p1 ? pos[p + stride] : make_float4(0.f, 0.f, 0.f, 0.f);
        n1 = pb + stride < p1 ? pos[pb + stride] : make_float4(0.f, 0.f, 0.f, 0.f);
        const DepositParticle q[2] = {deposit_particle(a, true, p, p0, P, lds_limit32),
                                      deposit_particle(b2, pb < p1, pb, p0, P, lds_limit32)};
        deposit_group<2, 1>(0, gbase, cells, P, q, frot, rot, off, hist, dst, defer_p, defer_g,
                            &defer_n);
      }
    } else {
      // the next particle is loaded while this one is deposited
      float4 qn = p0 + threadIdx.x < p1 ? pos[p0 + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
      for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
        const float4 qv = qn;
        qn = p + blockDim.x < p1 ? pos[p + blockDim.x] : make_float4(0.f, 0.f, 0.f, 0.f);
        const DepositParticle q[1] = {deposit_particle(qv, true, p, p0, P, lds_limit32)};
        int gi0 = 0;
        for (; gi0 + 4 <= ng; gi0 += 4)
          deposit_group<1, 4>(gi0, gbase, cells, P, q, frot, rot, off, hist, dst, defer_p, defer_g,
                              &defer_n);
        for (; gi0 < ng; ++gi0)
          deposit_group<1, 1>(gi0, gbase, cells, P, q, frot, rot, off, hist, dst, defer_p, defer_g,
                              &defer_n);
      }
    }
    __syncthreads();
    // the deferred pairs, one per lane
    const unsigned int nd = min(defer_n, (unsigned int)kBatchDeferCap);
    for (unsigned int d = threadIdx.x; d < nd; d += blockDim.x) {
      const int gi = defer_g[d];
      const int gr = gbase + gi;
      const float4 q = pos[p0 + defer_p[d]];
      const int cell = cell_exact_call(q.x - P.half, q.y - P.half, q.z - P.half, rot + 12 * gr,
                                       off + 4 * gr, P.W, P.half, P.h, P.lg);
      atomicAdd(&hist[gi * cells + cell], (unsigned int)(q.w * 1000.0f));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const unsigned int v = hist[i];
      if (v) atomicAdd(&dst[i], (unsigned long long)v);
    }
    __syncthreads();  // LDS, `cur` and the queue are reused by the next task
  }
}

hipError_t launch_batch_deposit(const float4* pos, const BatchSim* sims, const DepositTask* tasks,
                                int ntasks, const float* rot, const float* off, const float* frot,
                                unsigned long long* grids, unsigned long long* counter,
                                int threads, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  const size_t lds = (size_t)kBatchLdsCells * sizeof(unsigned int);
  // the request is the dynamic size itself: the kernel also has static LDS,
  // and static + dynamic may not exceed the 160 KB of a CU
  hipError_t e = hipFuncSetAttribute((const void*)batch_deposit_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // one workgroup per CU holds the 128 KB histogram budget
  const unsigned wg = (unsigned)(ntasks < 256 ? ntasks : 256);
  hipLaunchKernelGGL(batch_deposit_kernel, dim3(wg), dim3(threads), lds, s, pos, sims, tasks,
                     ntasks, rot, off, frot, grids, counter);
  return hipGetLastError();
}

// Blocks of kBatchGatherBlock particles, each inside one simulation, so the
// simulation (and with it every transform row) is uniform across the
// workgroup and comes through scalar loads, as in the single-simulation
// gather.
__global__ __launch_bounds__(kBatchGatherBlock) void batch_gather_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, long long blocks,
    const BatchSim* __restrict__ sims, int nsims, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    const float4* __restrict__ force, const int* __restrict__ order) {
  for (long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const BatchSim& b = sims[sim_of_block(sims, nsims, blk)];
    const long long j = (blk - b.blk_off) * kBatchGatherBlock + threadIdx.x;
    if (j >= b.n) continue;
    // the order permutes particles inside the simulation's own rows
    const long long i = order ? (long long)order[b.p_off + j] : b.p_off + j;
    const Params P = b.P;
    const float* R = rot + 12 * b.grid_off;
    const float* O = off + 4 * b.grid_off;
    const float* FR = frot + 12 * b.grid_off;
    const float4* F = force + b.cell_off;
    float4 p = pos[i];
    const float c0 = p.x - P.half, c1 = p.y - P.half, c2 = p.z - P.half;
    const int pfast = particle_fast(c0, c1, c2, P.CB);
    float tx = 0.f, ty = 0.f, tz = 0.f;
    // groups of U grids: the fast cells and their fetches first, then one
    // wave-uniform test re-fetches undecided pairs from the exact cells; the
    // forces are summed in grid order, as in the single-simulation gather
    constexpr int U = 4;
    int gr = 0;
    for (; gr + U <= P.k; gr += U) {
      int cell[U];
      bool ok[U];
      bool rare = false;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = cell_fast(c0, c1, c2, FR + 12 * (gr + u), P.lim, P.lg, &cell[u]) & pfast;
        rare |= !ok[u];
      }
      float4 f[U];
#pragma unroll
      for (int u = 0; u < U; ++u) f[u] = cell_load(F + (long long)(gr + u) * P.cells, cell[u]);
      if (__ballot(rare)) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!ok[u]) {
            const int ex = cell_exact_call(c0, c1, c2, R + 12 * (gr + u), O + 4 * (gr + u), P.W,
                                           P.half, P.h, P.lg);
            f[u] = cell_load(F + (long long)(gr + u) * P.cells, ex);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        tx += f[u].x;
        ty += f[u].y;
        tz += f[u].z;
      }
    }
    for (; gr < P.k; ++gr) {
      const int cell = cell_index(c0, c1, c2, pfast, FR + 12 * gr, R + 12 * gr, O + 4 * gr, P);
      const float4 f = cell_load(F + (long long)gr * P.cells, cell);
      tx += f.x;
      ty += f.y;
      tz += f.z;
    }
    float4 v = vel[i];
    integrate(p, v, tx / P.fk, ty / P.fk, tz / P.fk, P);
    pos[i] = p;
    vel[i] = v;
  }
}

hipError_t launch_batch_gather(float4* pos, float4* vel, long long blocks, const BatchSim* sims,
                               int nsims, const float* rot, const float* off, const float* frot,
                               const float4* force, const int* order, hipStream_t s) {
  if (blocks <= 0) return hipSuccess;
  const long long grid = blocks < 16384 ? blocks : 16384;
  hipLaunchKernelGGL(batch_gather_kernel, dim3((unsigned)grid), dim3(kBatchGatherBlock), 0, s, pos,
                     vel, blocks, sims, nsims, rot, off, frot, force, order);
  return hipGetLastError();
}

}  // namespace mog
