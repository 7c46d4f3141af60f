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
#include <type_traits>

#include "mog_common.h"

namespace mog {

// Repair records per lane and task (see batch_deposit_kernel).
constexpr int kBatchLaneSlots = 4;

// The 32-bit milli-mass a deposit adds for a particle of mass w (the
// single-simulation deposit's rule): uint(w * 1000) in [1, 2^32), clamped
// below 2^32 above, 0 below 1 and for NaN; `small` when it is the true
// milli-mass and fits the task's 32-bit partials.
__device__ __forceinline__ unsigned int batch_milli32(float w, unsigned int lds_limit32,
                                                      bool& small) {
  const float mm = w * 1000.0f;
  const unsigned int m = (unsigned int)fminf(fmaxf(mm, 0.0f), 4294967040.0f);
  small = m <= lds_limit32;
  return m;
}

// Persistent deposit over the task list.  A workgroup takes the next task
// from a device counter and deposits like the single-simulation kernel:
// every (particle, grid) pair adds its 32-bit milli-mass at the fast cell
// with no per-pair compare or select; a lane whose iteration holds an
// undecided pair (a face within the fast path's margin, a particle outside
// the admitted bound, a milli-mass that does not fit the 32-bit partials)
// records the iteration, and after the task's loop repairs its pairs: the
// exact cell takes the mass from the fast one (32-bit modular arithmetic),
// huge masses leave the partials for the 64-bit grids.  The flush adds the
// non-zero counters to the grids and zeroes them, so the next task starts on
// a clean histogram (one clear per workgroup, not one per task).
__global__ __launch_bounds__(1024) void batch_deposit_kernel(
    const float4* __restrict__ pos, const BatchSim* __restrict__ sims,
    const DepositTask* __restrict__ tasks, int ntasks, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    unsigned long long* __restrict__ grids, unsigned long long* __restrict__ counter) {
  extern __shared__ unsigned int hist[];  // kBatchLdsCells counters, then the records
  unsigned int* slots = hist + kBatchLdsCells;  // record r of lane t: slots[r * blockDim.x + t]
  __shared__ int cur;
  for (int i = threadIdx.x; i < kBatchLdsCells; i += blockDim.x) hist[i] = 0u;
  for (;;) {
    if (threadIdx.x == 0) cur = (int)atomicAdd(counter, 1ull);
    __syncthreads();
    const int ti = cur;
    if (ti >= ntasks) break;  // uniform across the workgroup
    const DepositTask task = tasks[ti];
    const BatchSim& b = sims[task.sim];
    const Params P = b.P;
    const int cells = P.cells;
    const int ng = task.ng;
    const int total = ng * cells;
    const long long p0 = b.p_off + task.p0, p1 = b.p_off + task.p1;
    const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)(task.p1 - task.p0);
    const unsigned int lds_limit32 =
        (unsigned int)(lds_limit < 0xFFFFFFFFull ? lds_limit : 0xFFFFFFFFull);
    const int gbase = b.grid_off + task.grid0;
    unsigned long long* dst = grids + b.cell_off + (long long)task.grid0 * cells;
    // one grid: two particles per lane and iteration; more: one particle
    const int UPn = ng == 1 ? 2 : 1;  // particles per lane and iteration
    const long long step = (long long)UPn * blockDim.x;

    // the pairs of the iteration whose first particle is p (particles p and,
    // with one grid, p + blockDim.x): add every fast cell when `add`, and
    // return whether every pair is decided
    auto iteration = [&](long long p, bool add, auto up) -> bool {
      constexpr int UP = decltype(up)::value;
      bool sure = true;
      float dmax = 0.0f;
      float4 qs[UP];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const long long pu = p + (long long)u * blockDim.x;
        qs[u] = pu < p1 ? pos[pu] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const float4 q = qs[u];
        bool small;
        const unsigned int m32 = batch_milli32(q.w, lds_limit32, small);
        const float c0 = q.x - P.half, c1 = q.y - P.half, c2 = q.z - P.half;
        sure = sure & small & (fabsf(c0) <= P.CB) & (fabsf(c1) <= P.CB) & (fabsf(c2) <= P.CB);
        int gi = 0;
        for (; gi + 4 <= ng; gi += 4) {
          int cell[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float* F = frot + 12 * (gbase + gi + v);
            dmax = fmaxf(dmax, cell_fast_d(c0, c1, c2, F, F[3], F[7], F[11], P.lg, &cell[v]));
          }
          if (add) {
#pragma unroll
            for (int v = 0; v < 4; ++v) atomicAdd(&hist[(gi + v) * cells + cell[v]], m32);
          }
        }
        for (; gi < ng; ++gi) {
          int cell;
          const float* F = frot + 12 * (gbase + gi);
          dmax = fmaxf(dmax, cell_fast_d(c0, c1, c2, F, F[3], F[7], F[11], P.lg, &cell));
          if (add) atomicAdd(&hist[gi * cells + cell], m32);
        }
      }
      return sure & (dmax < P.lim);
    };

    // repair the pairs of the iteration whose first particle is p
    auto repair = [&](long long p) {
      for (int u = 0; u < UPn; ++u) {
        const long long pu = p + (long long)u * blockDim.x;
        if (pu >= p1) continue;
        const float4 q = pos[pu];
        bool small;
        const unsigned int m32 = batch_milli32(q.w, lds_limit32, small);
        const float c0 = q.x - P.half, c1 = q.y - P.half, c2 = q.z - P.half;
        const int pf = particle_fast(c0, c1, c2, P.CB);
        for (int gi = 0; gi < ng; ++gi) {
          const int gr = gbase + gi;
          const float* F = frot + 12 * gr;
          int fc;
          const bool ok = pf && cell_fast_d(c0, c1, c2, F, F[3], F[7], F[11], P.lg, &fc) < P.lim;
          if (!small) {  // huge mass: out of the partials, into the 64-bit grids
            if (m32) atomicAdd(&hist[gi * cells + fc], 0u - m32);
            const unsigned long long milli = milli_of(q.w);
            if (milli) {
              const int ex = ok ? fc : cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr,
                                                       P.W, P.half, P.h, P.lg);
              atomicAdd(&dst[(long long)gi * cells + ex], milli);
            }
          } else if (!ok && m32) {
            const int ex = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half,
                                           P.h, P.lg);
            if (ex != fc) {
              atomicAdd(&hist[gi * cells + ex], m32);
              atomicAdd(&hist[gi * cells + fc], 0u - m32);
            }
          }
        }
      }
    };

    int nrec = 0;
    long long ovf = -1;
    auto record = [&](long long p) {
      if (nrec < kBatchLaneSlots)
        slots[nrec * blockDim.x + threadIdx.x] = (unsigned int)(p - p0);
      else if (nrec == kBatchLaneSlots)
        ovf = p;
      ++nrec;
    };
    // lane-varying trip counts: no barrier inside
    if (ng == 1) {
      for (long long p = p0 + threadIdx.x; p < p1; p += step)
        if (!iteration(p, true, std::integral_constant<int, 2>{})) record(p);
    } else {
      for (long long p = p0 + threadIdx.x; p < p1; p += step)
        if (!iteration(p, true, std::integral_constant<int, 1>{})) record(p);
    }
    const int mine = min(nrec, kBatchLaneSlots);
    for (int r = 0; r < mine; ++r) repair(p0 + slots[r * blockDim.x + threadIdx.x]);
    if (ovf >= 0) {  // more undecided iterations than records: find the rest again
      for (long long p = ovf; p < p1; p += step) {
        const bool sure = ng == 1 ? iteration(p, false, std::integral_constant<int, 2>{})
                                  : iteration(p, false, std::integral_constant<int, 1>{});
        if (!sure) repair(p);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const unsigned int v = hist[i];
      if (v) {
        atomicAdd(&dst[i], (unsigned long long)v);
        hist[i] = 0u;
      }
    }
    __syncthreads();  // LDS and `cur` are reused by the next task
  }
}

hipError_t launch_batch_deposit(const float4* pos, const BatchSim* sims, const DepositTask* tasks,
                                int ntasks, const float* rot, const float* off, const float* frot,
                                unsigned long long* grids, unsigned long long* counter,
                                int threads, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  const size_t lds = (size_t)kBatchLdsCells * sizeof(unsigned int) +
                     (size_t)kBatchLaneSlots * threads * sizeof(unsigned int);
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
// gather, whose deferral of undecided pairs (and so its summation order) it
// repeats.
__global__ __launch_bounds__(kBatchGatherBlock) void batch_gather_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, long long blocks,
    const BatchSim* __restrict__ sims, int nsims, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    const float4* __restrict__ force, const int* __restrict__ order) {
  __shared__ unsigned recs[kGatherSlots * kBatchGatherBlock];
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
    // wave-uniform test records undecided pairs, whose exact cells follow the
    // loop; the summation order is the single-simulation gather's
    constexpr int U = 4;
    int nrec = 0;
    int gr = 0;
    for (; gr + U <= P.k; gr += U) {
      // one compare per group: the largest face distance of its U pairs
      int cell[U];
      float d[U];
      float dmax = 0.0f;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float* Fr = FR + 12 * (gr + u);
        d[u] = cell_fast_d(c0, c1, c2, Fr, Fr[3], Fr[7], Fr[11], P.lg, &cell[u]);
        dmax = fmaxf(dmax, d[u]);
      }
      const bool rare = !(pfast & (dmax < P.lim));
      float4 f[U];
#pragma unroll
      for (int u = 0; u < U; ++u) f[u] = cell_load(F + (long long)(gr + u) * P.cells, cell[u]);
      if (__ballot(rare)) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!(pfast & (d[u] < P.lim))) {
            if (nrec < kGatherSlots) {
              recs[nrec * kBatchGatherBlock + threadIdx.x] = (unsigned)(gr + u);
              f[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            } else {
              const int ex = cell_exact_call(c0, c1, c2, R + 12 * (gr + u), O + 4 * (gr + u), P.W,
                                             P.half, P.h, P.lg);
              f[u] = cell_load(F + (long long)(gr + u) * P.cells, ex);
            }
            ++nrec;
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
    const int mine = min(nrec, kGatherSlots);
    for (int r = 0; r < mine; ++r) {
      const int gd = (int)recs[r * kBatchGatherBlock + threadIdx.x];
      const int ex =
          cell_exact_call(c0, c1, c2, R + 12 * gd, O + 4 * gd, P.W, P.half, P.h, P.lg);
      const float4 f = cell_load(F + (long long)gd * P.cells, ex);
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
