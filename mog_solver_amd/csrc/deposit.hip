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
#include <type_traits>

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

typedef __attribute__((address_space(3))) unsigned int lds_u32;

// Deferred-pair records per lane of the LDS deposit (see deposit_lds_kernel).
constexpr int kLaneSlots = 4;

// UP particles per lane and register set in the LDS deposit.
constexpr int kDepositUP = 2;

// One tile of the LDS-privatised deposit: particles [p0, p1) of one
// simulation (rows of pos / cpos) into its grids [grid0, grid0 + ng), ng <=
// G, through G uint32 histograms at `hist` (zero on entry; at LDS address 0
// under SH) and kLaneSlots records per lane at `slots`.  Called by every
// thread of the workgroup (it holds barriers).  The G fast rows are
// wave-uniform and live in scalar (for G <= 4 under SH: vector) registers
// for the whole tile.  Used by deposit_lds_kernel (one tile per workgroup)
// and by the batched ensemble's persistent deposit (one tile per task).
//
// Every (particle, grid) pair adds the particle's 32-bit milli-mass at its
// fast cell (cellmath.h), unconditionally: no per-pair compare or select.
// A lane whose UP x G pairs of one iteration include one the fast path
// cannot decide (a face within the error margin, a particle outside the
// admitted bound or non-finite, a milli-mass that does not fit the 32-bit
// partials) appends one record, the offset of its first particle, to its
// own LDS slots; after the loop each lane repairs its recorded pairs: the
// exact float32 sequence gives the cell, and if it differs the mass moves
// (added at the exact cell, subtracted at the fast one: 32-bit modular
// arithmetic, exact since the chunk's final sums fit), and huge masses are
// taken out of the partials and added to the 64-bit grids in HBM.  A lane
// with more records than slots repairs on the spot.  The counters are exact
// while a chunk's sum stays below 2^32: a particle whose milli-mass exceeds
// (2^32-1)/chunk takes the 64-bit path.
//
// Two loops feed the same per-particle work.  In a tile whose group holds G
// live grids, the whole (2 x UP x threads)-particle spans of its range run a
// double-buffered loop: two register sets of UP particles each, one
// processed while the other loads, with unguarded loads at 32-bit byte
// offsets from the range's base.  The rest run the guarded loop.
//
// SH: the fast cells come as LDS byte offsets through the shifted fields of
// cellmath.h (cell_fast_o), for launches whose transforms satisfy
// field_shift_ok(lg, 2, ...); otherwise as indices (cell_fast_d).
//
// PREP: the loops read the rows deposit_prep_kernel wrote (centred
// coordinates and the 32-bit milli-mass, or kPrepMarker for a particle the
// fast path cannot take: outside the admitted bound, non-finite or with a
// milli-mass that does not fit the partials), so the per-particle work of a
// pair is one compare; the repair adds the marker back and takes such
// particles from `pos`.
template <int G, bool SH, bool PREP>
__device__ __forceinline__ void deposit_tile(
    unsigned int* __restrict__ hist, unsigned int* __restrict__ slots,
    const float4* __restrict__ pos, const float4* __restrict__ cpos, long long p0, long long p1,
    int grid0, int ng, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot, const Params& P,
    unsigned long long lds_limit, unsigned long long* __restrict__ grids, bool rezero) {
  const int cells = P.cells;
  const int total = ng * cells;
  float F[G][12];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int gr = min(grid0 + gi, P.k - 1);
#pragma unroll
    for (int j = 0; j < 12; ++j) F[gi][j] = frot[12 * gr + j];
    if constexpr (SH && G <= 4) {  // all 12 in vector registers: full-rate FMAs
#pragma unroll
      for (int j = 0; j < 12; ++j) F[gi][j] = to_vgpr(F[gi][j]);
    } else {
#pragma unroll
      for (int m = 0; m < 3; ++m) F[gi][4 * m + 3] = to_vgpr(F[gi][4 * m + 3]);
    }
  }
  __syncthreads();

  const unsigned int lds_limit32 =
      (unsigned int)(lds_limit < 0xFFFFFFFFull ? lds_limit : 0xFFFFFFFFull);
  constexpr int UP = kDepositUP;

  // The 32-bit milli-mass the loop adds for a particle: uint(m * 1000) when
  // the product is in [1, 2^32), 2^32 - 1 from 2^32 on, 0 below 1 and for
  // NaN (the saturating conversion of v_cvt_u32_f32); `small` when it is the
  // true milli-mass and fits the partials (lds_limit32 <= (2^32 - 1) / 256,
  // so a saturated value is never small).
  auto milli32 = [&](float w, bool& small) -> unsigned int {
    const float mm = w * 1000.0f;
    unsigned int m;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(m) : "v"(mm));
    small = m <= lds_limit32;
    return m;
  };
  FieldShift FS = field_shift(P.lg, 2);
  float half = P.half;
  if constexpr (SH) {  // vector-register operands issue at the full rate
    FS.k1 = to_vgpr(FS.k1);
    FS.k2 = to_vgpr(FS.k2);
    FS.m0 = to_vgpr(FS.m0);
    FS.m01 = to_vgpr(FS.m01);
    FS.m012 = to_vgpr(FS.m012);
    half = to_vgpr(half);
  }
  // pair adds: (grid gi, fast cell c) += v; under SH, c is the LDS address
  // of the counter itself (hist starts at LDS address 0: the kernel holds
  // no static LDS, which the launch checks)
  auto add_at = [&](int gi, unsigned c, unsigned v) {
    if constexpr (SH)
      __hip_atomic_fetch_add((lds_u32*)(size_t)c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      atomicAdd(hist + gi * cells + c, v);
  };

  // The fast cell (index) of a pair and its face distance, computed as the
  // loop computed the cell it added to: the two forms agree for admitted
  // particles only, and a repair takes back what was added.
  auto fast_cell = [&](float c0, float c1, float c2, int gi, int* fc) -> float {
    if constexpr (SH) {
      unsigned o;
      const float d = cell_fast_o(c0, c1, c2, F[gi], F[gi][3], F[gi][7], F[gi][11], FS, 0u, &o);
      *fc = (int)(o >> 2);
      return d;
    } else {
      return cell_fast_d(c0, c1, c2, F[gi], F[gi][3], F[gi][7], F[gi][11], P.lg, fc);
    }
  };

  // PREP: the pairs of particle p.  A marked particle had the marker added
  // at its fast cells (2^32 - 1, i.e. -1 modulo 2^32): it is taken back and
  // the true milli-mass goes to the exact cell, in the partials when small,
  // else into the 64-bit grids.
  auto repair_prepped = [&](long long p) {
    const float4 cp = cpos[p];
    const unsigned int m_add = __float_as_uint(cp.w);
    const float c0 = cp.x, c1 = cp.y, c2 = cp.z;
    const int pf = particle_fast(c0, c1, c2, P.CB);
    const bool marked = m_add > lds_limit32;
    bool small = true;
    unsigned int m32 = 0;
    float w = 0.0f;
    if (marked) {
      w = pos[p].w;
      m32 = milli32(w, small);
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      if (gi >= ng) continue;
      const int gr = grid0 + gi;
      int fc;
      const float d = fast_cell(c0, c1, c2, gi, &fc);
      const bool ok = pf && d < P.lim;
      if (marked) {
        atomicAdd(&hist[gi * cells + fc], 0u - m_add);
        const unsigned long long milli = small ? m32 : milli_of(w);
        if (milli) {
          const int ex = ok ? fc : cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W,
                                                   P.half, P.h, P.lg);
          if (small)
            atomicAdd(&hist[gi * cells + ex], m32);
          else
            atomicAdd(&grids[(long long)gr * cells + ex], milli);
        }
      } else if (!ok && m_add) {
        const int ex = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half, P.h,
                                       P.lg);
        if (ex != fc) {
          atomicAdd(&hist[gi * cells + ex], m_add);
          atomicAdd(&hist[gi * cells + fc], 0u - m_add);
        }
      }
    }
  };

  // Repair the pairs of the UP particles of one record (first particle
  // pu0): exact cell where the fast one was not decided, 64-bit path for
  // huge masses.  Runs per lane.
  auto repair = [&](long long pu0) {
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      const long long p = pu0 + (long long)u * blockDim.x;
      if (p >= p1) continue;
      if constexpr (PREP) {
        repair_prepped(p);
        continue;
      }
      const float4 q = pos[p];
      bool small;
      const unsigned int m32 = milli32(q.w, small);
      const float c0 = q.x - P.half, c1 = q.y - P.half, c2 = q.z - P.half;
      const int pf = particle_fast(c0, c1, c2, P.CB);
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
        if (gi >= ng) continue;
        const int gr = grid0 + gi;
        int fc;
        const float d = fast_cell(c0, c1, c2, gi, &fc);
        const bool ok = pf && d < P.lim;
        if (!small) {  // huge mass: out of the partials, into the 64-bit grids
          if (m32) atomicAdd(&hist[gi * cells + fc], 0u - m32);
          const unsigned long long milli = milli_of(q.w);
          if (milli) {
            const int ex = ok ? fc : cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W,
                                                     P.half, P.h, P.lg);
            atomicAdd(&grids[(long long)gr * cells + ex], milli);
          }
        } else if (!ok && m32) {
          const int ex = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half, P.h,
                                         P.lg);
          if (ex != fc) {
            atomicAdd(&hist[gi * cells + ex], m32);
            atomicAdd(&hist[gi * cells + fc], 0u - m32);
          }
        }
      }
    }
  };

  // The UP x G pairs of UP particles: their milli-masses and fast cells,
  // every cell computed before any is used (no branch between them); true
  // when every pair is decided.
  auto cells_of = [&](const float4* q, unsigned int (&m32)[UP], unsigned (&cell)[UP][G]) -> bool {
    float cc[UP][3];
    bool sure = true;  // flags combine with bitwise &: no short-circuit branches
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      if constexpr (PREP) {  // admission and mass folded into one compare
        m32[u] = __float_as_uint(q[u].w);
        cc[u][0] = q[u].x;
        cc[u][1] = q[u].y;
        cc[u][2] = q[u].z;
        sure = sure & (m32[u] <= lds_limit32);
      } else {
        bool small;
        m32[u] = milli32(q[u].w, small);
        cc[u][0] = q[u].x - half;
        cc[u][1] = q[u].y - half;
        cc[u][2] = q[u].z - half;
        sure = sure & small & (fabsf(cc[u][0]) <= P.CB) & (fabsf(cc[u][1]) <= P.CB) &
               (fabsf(cc[u][2]) <= P.CB);
      }
    }
    float dmax = 0.0f;
#pragma unroll
    for (int u = 0; u < UP; ++u)
#pragma unroll
      for (int gi = 0; gi < G; ++gi) {
        if constexpr (SH) {
          dmax = fmaxf(dmax, cell_fast_o(cc[u][0], cc[u][1], cc[u][2], F[gi], F[gi][3], F[gi][7],
                                         F[gi][11], FS, (unsigned)(gi * cells * 4), &cell[u][gi]));
        } else {
          int c;
          dmax = fmaxf(dmax, cell_fast_d(cc[u][0], cc[u][1], cc[u][2], F[gi], F[gi][3], F[gi][7],
                                         F[gi][11], P.lg, &c));
          cell[u][gi] = (unsigned)c;
        }
      }
    return sure & (dmax < P.lim);
  };

  // A lane's iterations start at chunk offsets threadIdx.x + j * UP *
  // threads in both loops.  Records: the first kLaneSlots undecided
  // iterations; from the next one on (ovf) the lane rescans after the loop.
  int nrec = 0;
  long long ovf = -1;
  auto pairs = [&](const float4* q, long long pl) {
    unsigned int m32[UP];
    unsigned cell[UP][G];
    const bool sure = cells_of(q, m32, cell);
    // grids gi >= ng (last group) add to allocated, unflushed counters
#pragma unroll
    for (int u = 0; u < UP; ++u)
#pragma unroll
      for (int gi = 0; gi < G; ++gi) add_at(gi, cell[u][gi], m32[u]);
    if (!sure) {  // rare
      if (nrec < kLaneSlots)
        slots[nrec * blockDim.x + threadIdx.x] = (unsigned int)pl;
      else if (nrec == kLaneSlots)
        ovf = pl;
      ++nrec;
    }
  };

  // whole spans of the range (when the group holds G live grids) take the
  // double-buffered loop, the rest the guarded one
  const long long span = 2ll * UP * blockDim.x;
  const long long spans = (ng == G && p1 - p0 < (1ll << 27)) ? (p1 - p0) / span : 0;
  const float4* const src = PREP ? cpos : pos;  // the rows the loops read
  if (spans > 0) {
    const char* base = reinterpret_cast<const char*>(src + p0);
    const unsigned int lane_step = 16u * blockDim.x;  // bytes between a lane's particles
    auto load = [&](float4* q, unsigned int boff) {
#pragma unroll
      for (int u = 0; u < UP; ++u)
        q[u] = *reinterpret_cast<const float4*>(base + boff + (unsigned int)u * lane_step);
    };
    const int iters = (int)spans;
    unsigned int boff = 16u * threadIdx.x;
    float4 a[UP], b[UP];
    load(a, boff);
    for (int it = 0; it < iters; ++it) {
      const unsigned int boff_b = boff + UP * lane_step;
      load(b, boff_b);
      pairs(a, boff >> 4);
      boff = boff_b + UP * lane_step;
      if (it + 1 < iters) load(a, boff);
      pairs(b, boff_b >> 4);
    }
  }
  if (p0 + spans * span < p1) {
    const long long step = (long long)UP * blockDim.x;
    const long long pfirst = p0 + spans * span + threadIdx.x;
    // the next iteration's particles are loaded while this one computes;
    // rows past the range read as mass 0 (they add 0 and are never repaired)
    float4 qn[UP];
#pragma unroll
    for (int u = 0; u < UP; ++u) {
      const long long p = pfirst + (long long)u * blockDim.x;
      qn[u] = p < p1 ? src[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    for (long long pb = pfirst; pb < p1; pb += step) {
      float4 q[UP];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        q[u] = qn[u];
        const long long p = pb + step + (long long)u * blockDim.x;
        qn[u] = p < p1 ? src[p] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      pairs(q, pb - p0);
    }
  }
  // the recorded pairs, per lane
  const int mine = min(nrec, kLaneSlots);
  for (int r = 0; r < mine; ++r) repair(p0 + slots[r * blockDim.x + threadIdx.x]);
  if (ovf >= 0) {  // more undecided iterations than slots: find the rest again
    for (long long pl = ovf; p0 + pl < p1; pl += (long long)UP * blockDim.x) {
      float4 q[UP];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const long long p = p0 + pl + (long long)u * blockDim.x;
        q[u] = p < p1 ? src[p] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      unsigned int m32[UP];
      unsigned cell[UP][G];
      if (!cells_of(q, m32, cell)) repair(p0 + pl);
    }
  }
  __syncthreads();
  // flush the live counters; with `rezero` every counter the tile touched
  // (its G histograms, live or not) is left at zero for the next tile
  unsigned long long* dst = grids + (long long)grid0 * cells;
  for (int i = threadIdx.x; i < (rezero ? G * cells : total); i += blockDim.x) {
    const unsigned int v = hist[i];
    if (v) {
      if (i < total) atomicAdd(&dst[i], (unsigned long long)v);
      if (rezero) hist[i] = 0u;
    }
  }
}

// LDS-privatised deposit; grid (chunks, ceil(k / G)) in XCD-aware,
// chunk-major order.  Dynamic LDS holds G * cells uint32 counters followed
// by kLaneSlots records per lane (and no static LDS: the shifted-field path
// addresses counters by their LDS address from 0).
template <int G, bool SH, bool PREP>
__global__ __launch_bounds__(1024) void deposit_lds_kernel(
    const float4* __restrict__ pos, const float4* __restrict__ cpos, long long n, long long chunk,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    Params P, unsigned long long lds_limit, unsigned long long* __restrict__ grids) {
  extern __shared__ unsigned int hist[];
  // tiles in chunk-major order: the groups of one chunk share an XCD's L2
  const int groups = (P.k + G - 1) / G;
  const unsigned int tile = xcd_tile(blockIdx.x, gridDim.x);
  const long long chunk_id = tile / groups;
  const int grid0 = (int)(tile % groups) * G;
  const int ng = min(G, P.k - grid0);
  for (int i = threadIdx.x; i < ng * P.cells; i += blockDim.x) hist[i] = 0u;
  const long long p0 = chunk_id * chunk;
  deposit_tile<G, SH, PREP>(hist, hist + G * P.cells, pos, cpos, p0, min(n, p0 + chunk), grid0,
                            ng, rot, off, frot, P, lds_limit, grids, false);
}

// Grids per task of the batched deposit: the largest power of two up to
// kBatchMaxG whose histograms fit kBatchLdsCells (G <= 4 keeps the fast rows
// in vector registers).
int batch_task_grids(int cells) {
  int G = 1;
  while (G * 2 <= kBatchMaxG && (G * 2) * cells <= kBatchLdsCells) G *= 2;
  return G;
}

// Prepared rows of every particle of the batch (the rows of
// deposit_prep_kernel with each simulation's own centre and admission
// bound), one workgroup of kBatchGatherBlock threads per gather block (a
// block lies inside one simulation).  lds_limit32: the partials' limit of
// the batch's largest task, (2^32 - 1) / max(deposit chunk, 256); a task
// of fewer particles has a higher limit, so the rows a task finds unmarked
// fit its partials.
__global__ __launch_bounds__(kBatchGatherBlock) void batch_prep_kernel(
    const float4* __restrict__ pos, const BatchSim* __restrict__ sims, int nsims,
    unsigned int lds_limit32, float4* __restrict__ cpos) {
  const BatchSim& b = sims[sim_of_block(sims, nsims, blockIdx.x)];
  const long long j = (blockIdx.x - b.blk_off) * kBatchGatherBlock + threadIdx.x;
  if (j >= b.n) return;
  const long long i = b.p_off + j;
  cpos[i] = prepared_row(pos[i], b.P, lds_limit32);
}

// Persistent deposit of the batched ensemble over its task list (particle
// chunk x group of grids of one simulation, ordered by cost, largest first:
// dynamic pickup balances simulations whose n spans 50x, k 7.5x, g^3 8x).
// A workgroup takes the next task from a device counter and runs it as one
// deposit_tile of the single-simulation deposit, so the grids are the same
// integers; the tile's flush leaves its counters at zero for the next task
// (one clear per workgroup, not one per task).  Dynamic LDS: kBatchLdsCells
// counters from LDS address 0, the lanes' records, then the task index.
__global__ __launch_bounds__(1024) void batch_deposit_kernel(
    const float4* __restrict__ pos, const float4* __restrict__ cpos,
    const BatchSim* __restrict__ sims, const DepositTask* __restrict__ tasks, int ntasks,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    unsigned long long* __restrict__ grids, unsigned long long* __restrict__ counter) {
  extern __shared__ unsigned int hist[];
  unsigned int* const slots = hist + kBatchLdsCells;
  int* const cur = reinterpret_cast<int*>(slots + kLaneSlots * blockDim.x);
  for (int i = threadIdx.x; i < kBatchLdsCells; i += blockDim.x) hist[i] = 0u;
  for (;;) {
    if (threadIdx.x == 0) *cur = (int)atomicAdd(counter, 1ull);
    __syncthreads();
    const int ti = *cur;
    if (ti >= ntasks) break;  // uniform across the workgroup
    const DepositTask task = tasks[ti];
    const BatchSim& b = sims[task.sim];
    const Params P = b.P;
    const long long len = task.p1 - task.p0;
    // 32-bit partials exact for the task's particles (at least 256 in the
    // divisor: a saturated milli-mass is then never taken as small)
    const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)(len > 256 ? len : 256);
    const float4* sp = pos + b.p_off;
    const float4* cp = cpos + b.p_off;
    const float* R = rot + 12 * b.grid_off;
    const float* O = off + 4 * b.grid_off;
    const float* FR = frot + 12 * b.grid_off;
    unsigned long long* gd = grids + b.cell_off;
#define MOG_TILE(GG, SHB)                                                                     \
  deposit_tile<GG, SHB, true>(hist, slots, sp, cp, task.p0, task.p1, task.grid0, task.ng, R,  \
                              O, FR, P, lds_limit, gd, true)
    if (P.shift_ok & kShiftLds) {
      if (task.G == 4) MOG_TILE(4, true);
      else if (task.G == 2) MOG_TILE(2, true);
      else MOG_TILE(1, true);
    } else {
      if (task.G == 4) MOG_TILE(4, false);
      else if (task.G == 2) MOG_TILE(2, false);
      else MOG_TILE(1, false);
    }
#undef MOG_TILE
    __syncthreads();  // LDS and the task index are reused by the next task
  }
}

hipError_t launch_batch_deposit(const float4* pos, float4* cpos, const BatchSim* sims, int nsims,
                                long long blocks, long long chunk, const DepositTask* tasks,
                                int ntasks, const float* rot, const float* off, const float* frot,
                                unsigned long long* grids, unsigned long long* counter,
                                int threads, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  const unsigned int lds_limit32 =
      (unsigned int)(0xFFFFFFFFull / (unsigned long long)(chunk > 256 ? chunk : 256));
  hipLaunchKernelGGL(batch_prep_kernel, dim3((unsigned)blocks), dim3(kBatchGatherBlock), 0, s, pos,
                     sims, nsims, lds_limit32, cpos);
  hipError_t pe = hipGetLastError();
  if (pe != hipSuccess) return pe;
  const size_t lds = (size_t)kBatchLdsCells * sizeof(unsigned int) +
                     (size_t)kLaneSlots * threads * sizeof(unsigned int) + 16;
  hipError_t e = hipFuncSetAttribute((const void*)batch_deposit_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  // one workgroup per CU holds the 128 KB histogram budget
  const unsigned wg = (unsigned)(ntasks < 256 ? ntasks : 256);
  hipLaunchKernelGGL(batch_deposit_kernel, dim3(wg), dim3(threads), lds, s, pos,
                     (const float4*)cpos, sims, tasks, ntasks, rot, off, frot, grids, counter);
  return hipGetLastError();
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
// pair is stored in cells_out[cache_index(t, grid, k)] for the gather, with
// bit 15 set where the fast path could not decide (the gather defers those
// pairs; g <= 32, so cells fit 15 bits).
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
        int cell;
        const int ok = cell_fast(c0, c1, c2, F[gi], P.lim, P.lg, &cell) & pfast;
        if (!ok) cell = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half, P.h, P.lg);
        if (CACHE && live)
          cells_out[cache_index(p, gr, P.k)] = (unsigned short)(cell | (ok ? 0 : 0x8000));
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

__global__ __launch_bounds__(256) void deposit_prep_kernel(const float4* __restrict__ pos,
                                                           long long n, Params P,
                                                           unsigned int lds_limit32,
                                                           float4* __restrict__ cpos) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cpos[i] = prepared_row(pos[i], P, lds_limit32);
}

template <int G>
static hipError_t launch_lds(const Params& P, const float4* pos, long long n, long long chunks,
                             long long chunk, int threads, const float* rot, const float* off,
                             const float* frot, unsigned long long* grids,
                             unsigned short* cells_out, bool sorted, float4* cpos, hipStream_t s,
                             bool cpos_ready, unsigned int* prep_limit) {
  const int groups = (P.k + G - 1) / G;
  const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)chunk;
  const size_t lds = (size_t)G * P.cells * 4 + (sorted ? 0 : (size_t)kLaneSlots * threads * 4);
  bool sh = P.shift_ok & kShiftLds;
  if (sh && !sorted) {  // the shifted path addresses its counters from LDS address 0
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, (const void*)deposit_lds_kernel<G, true, false>);
    if (e != hipSuccess) return e;
    sh = fa.sharedSizeBytes == 0;
    if constexpr (G <= 4) {
      if (sh && cpos) {
        e = hipFuncGetAttributes(&fa, (const void*)deposit_lds_kernel<G, true, true>);
        if (e != hipSuccess) return e;
        if (fa.sharedSizeBytes != 0) cpos = nullptr;
      }
    }
  }
  const bool prep = G <= 4 && sh && !sorted && cpos;
  const void* fn = sorted ? (cells_out ? (const void*)deposit_sorted_kernel<G, true>
                                       : (const void*)deposit_sorted_kernel<G, false>)
                   : sh   ? (const void*)deposit_lds_kernel<G, true, false>
                          : (const void*)deposit_lds_kernel<G, false, false>;
  if constexpr (G <= 4)
    if (prep) fn = (const void*)deposit_lds_kernel<G, true, true>;
  // the exact dynamic size: the kernels also hold static LDS
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)(chunks * groups));
  if constexpr (G <= 4) {
    if (prep) {
      const unsigned int lds_limit32 = (unsigned int)lds_limit;  // (2^32 - 1) / chunk
      if (prep_limit) *prep_limit = lds_limit32;
      if (!cpos_ready) {  // else the rows are already in cpos (written for this step)
        hipLaunchKernelGGL(deposit_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           s, pos, n, P, lds_limit32, cpos);
        if ((e = hipGetLastError()) != hipSuccess) return e;
      }
      hipLaunchKernelGGL((deposit_lds_kernel<G, true, true>), grid, dim3(threads), lds, s, pos,
                         (const float4*)cpos, n, chunk, rot, off, frot, P, lds_limit, grids);
      return hipGetLastError();
    }
  }
  if (!sorted && sh) {
    hipLaunchKernelGGL((deposit_lds_kernel<G, true, false>), grid, dim3(threads), lds, s, pos,
                       (const float4*)nullptr, n, chunk, rot, off, frot, P, lds_limit, grids);
  } else if (!sorted) {
    hipLaunchKernelGGL((deposit_lds_kernel<G, false, false>), grid, dim3(threads), lds, s, pos,
                       (const float4*)nullptr, n, chunk, rot, off, frot, P, lds_limit, grids);
  } else if (cells_out) {
    hipLaunchKernelGGL((deposit_sorted_kernel<G, true>), grid, dim3(threads), lds, s, pos, n, chunk,
                       rot, off, frot, P, lds_limit, grids, cells_out);
  } else {
    hipLaunchKernelGGL((deposit_sorted_kernel<G, false>), grid, dim3(threads), lds, s, pos, n,
                       chunk, rot, off, frot, P, lds_limit, grids, cells_out);
  }
  return hipGetLastError();
}

hipError_t launch_deposit_prep(const Params& P, const float4* pos, long long n,
                               unsigned int lds_limit32, float4* cpos, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(deposit_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                     n, P, lds_limit32, cpos);
  return hipGetLastError();
}

hipError_t launch_deposit(const Params& P, const float4* pos, long long n, const float* rot,
                          const float* off, const float* frot, unsigned long long* grids,
                          const DepositTune& t, hipStream_t s, float4* cpos,
                          unsigned int* prep_limit) {
  return launch_deposit_ex(P, pos, n, rot, off, frot, grids, nullptr, false, t, s, cpos,
                           prep_limit);
}

hipError_t launch_deposit_ex(const Params& P, const float4* pos, long long n, const float* rot,
                             const float* off, const float* frot, unsigned long long* grids,
                             unsigned short* cells_out, bool sorted, const DepositTune& t,
                             hipStream_t s, float4* cpos, unsigned int* prep_limit) {
  if (prep_limit) *prep_limit = 0;
  if (n <= 0) return hipSuccess;
  // by default the prepass runs from 2^26 pairs, where it pays for itself
  // (two disks, 2.4e8 pairs: deposit phase -1 %; 64 Plummer spheres,
  // 9.6e8: -9 %; 65,536 particles x 16 grids: +8 %, the extra launch and
  // pass cost more than the loop saves)
  if (t.prepass < 0 || (t.prepass == 0 && (long long)n * P.k < (1ll << 26))) cpos = nullptr;
  const size_t grid_bytes = (size_t)P.cells * 4;
  if (grid_bytes + (size_t)kLaneSlots * 1024 * 4 > kMaxLds) {
    const int G = 8;
    const int groups = (P.k + G - 1) / G;
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(deposit_global_kernel, dim3((unsigned)blocks, (unsigned)groups), dim3(256),
                       0, s, pos, n, G, rot, off, frot, P, grids);
    return hipGetLastError();
  }
  // G in {1, 2, 4, 8}, within the LDS and at most k
  int G = 1;
  // (the LDS also holds the per-lane repair records: at most 16 KB)
  while (G * 2 <= t.grids_per_wg && G * 2 <= 8 &&
         (size_t)(G * 2) * grid_bytes + (size_t)kLaneSlots * 1024 * 4 <= kMaxLds &&
         G * 2 <= P.k)
    G *= 2;
  const int groups = (P.k + G - 1) / G;
  const int threads = t.threads;
  long long chunks = (t.wg_target + groups - 1) / groups;
  // one-dimensional launch of chunks x groups workgroups: under 2^32 lanes
  const long long max_chunks = 0xFFFFFFFFll / ((long long)threads * groups);
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  // whole spans of 2 x 2 x threads particles: every chunk but the last then
  // takes the double-buffered, unguarded loop of deposit_lds_kernel
  const long long span = 4ll * threads;
  long long chunk = (n + chunks - 1) / chunks;
  if (chunk < t.min_chunk) chunk = t.min_chunk;
  chunk = (chunk + span - 1) / span * span;
  chunks = (n + chunk - 1) / chunk;
  switch (G) {
    case 8:
      return launch_lds<8>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, cpos, s, t.cpos_ready, prep_limit);
    case 4:
      return launch_lds<4>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, cpos, s, t.cpos_ready, prep_limit);
    case 2:
      return launch_lds<2>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, cpos, s, t.cpos_ready, prep_limit);
    default:
      return launch_lds<1>(P, pos, n, chunks, chunk, threads, rot, off, frot, grids, cells_out,
                           sorted, cpos, s, t.cpos_ready, prep_limit);
  }
}

}  // namespace mog
