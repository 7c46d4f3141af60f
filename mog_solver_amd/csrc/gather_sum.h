// The per-particle grid loop of the force gather (update.glsl:67-80),
// shared by the single-simulation gather (gather.hip) and the batched
// ensemble's gather (batch.hip), so that both sum the same forces in the
// same order.
#pragma once

#include "mog_common.h"

namespace mog {

// t += M F for a grid-frame force F of grid R: world_m = sum_j R[4m+j] F_j
// (transpose(rot) * gridForce, update.glsl:77).
__device__ __forceinline__ void accumulate_rotated(const float* __restrict__ R, const float4 F,
                                                   float& tx, float& ty, float& tz) {
  tx += fmaf(R[0], F.x, fmaf(R[1], F.y, R[2] * F.z));
  ty += fmaf(R[4], F.x, fmaf(R[5], F.y, R[6] * F.z));
  tz += fmaf(R[8], F.x, fmaf(R[9], F.y, R[10] * F.z));
}

// The force sum of one particle (centred position c0..c2, pfast: admitted
// to the fast path) over the P.k grids of one simulation: t += M F of the
// particle's NGP cell in every grid (update.glsl:67-80).
//
// WORLD: the force grids already hold M F (the step's solve rotates each
// cell once), so a pair costs three adds instead of nine FMAs.  The grid
// loop is unrolled by U so that U force fetches are in flight per wave.
//
// LDS3: the fast rows' addends staged in LDS, adds[3 g + m] =
// frot[12 g + 4 m + 3], so the FMA chain reads them as vector operands
// (ds_read) instead of re-copying them from scalar registers with three
// v_mov per pair; otherwise they come through the scalar cache and `adds`
// is not read.  (adds must point at the workgroup's LDS: the caller passes
// its __shared__ array itself, so that the reads compile to LDS loads.)
//
// SH: the fast cells come as byte offsets into the float4 grid through the
// shifted fields of cellmath.h (cell_fast_o), for transforms that satisfy
// field_shift_ok(lg, 4, ...); otherwise as indices (cell_fast_d).
//
// Undecided pairs (a face within the fast path's margin, or a particle
// outside the admitted bound) are deferred: inside the rarely taken branch a
// lane only records the grid in its LDS slots (recs[r * nth + lane]) and
// drops the fast fetch; after the grid loop it evaluates the exact cells of
// its records and adds their forces.  A wave then runs about one exact
// evaluation per particle pass (the most records of any of its lanes)
// instead of one per group in which any lane needs one.  The forces are
// summed in grid order except the deferred ones, which come last.
template <bool WORLD, bool LDS3, bool SH>
__device__ __forceinline__ void gather_sum(float c0, float c1, float c2, int pfast,
                                           const Params& P, const float* __restrict__ rot,
                                           const float* __restrict__ off,
                                           const float* __restrict__ frot,
                                           const float4* __restrict__ force, const float* adds,
                                           unsigned* recs, int nth, int lane, float& tx,
                                           float& ty, float& tz) {
  constexpr int U = 4;
  FieldShift FS = field_shift(P.lg, 4);
  if constexpr (SH) {  // vector-register operands issue at the full rate
    FS.k0 = to_vgpr(FS.k0);
    FS.k1 = to_vgpr(FS.k1);
    FS.k2 = to_vgpr(FS.k2);
    FS.m0 = to_vgpr(FS.m0);
    FS.m01 = to_vgpr(FS.m01);
    FS.m012 = to_vgpr(FS.m012);
  }
  // a wave with a particle outside the fast path takes the branch below for
  // every group (its lanes decide per pair); one wave-uniform flag, taken
  // once, so that the per-group test is a compare and a scalar branch
  const bool wave_slow = __builtin_amdgcn_ballot_w64(!pfast) != 0;
  int nrec = 0;  // undecided pairs met so far
  int grid = 0;
  for (; grid + U <= P.k; grid += U) {
    // the U fast cells (always in range) and their force fetches first, then
    // one wave-uniform test: only a wave with an undecided pair enters the
    // branch, which records those pairs (the exact sequence is a call, whose
    // entry waits for every load in flight: it runs after the loop)
    // one compare per group: the largest face distance of its U pairs
    unsigned cell[U];  // byte offsets under SH
    float d[U];
    float dmax = 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int gr = grid + u;
      const float* Fr = frot + 12 * gr;
      const float a0 = LDS3 ? adds[3 * gr] : Fr[3];
      const float a1 = LDS3 ? adds[3 * gr + 1] : Fr[7];
      const float a2 = LDS3 ? adds[3 * gr + 2] : Fr[11];
      if constexpr (SH) {
        d[u] = cell_fast_o(c0, c1, c2, Fr, a0, a1, a2, FS, 0u, &cell[u]);
      } else {
        int c;
        d[u] = cell_fast_d(c0, c1, c2, Fr, a0, a1, a2, P.lg, &c);
        cell[u] = (unsigned)c << 4;
      }
      dmax = fmaxf(dmax, d[u]);
    }
    float4 F[U];
#pragma unroll
    for (int u = 0; u < U; ++u) F[u] = cell_load_b(force + (long long)(grid + u) * P.cells, cell[u]);
    if (wave_slow || __builtin_amdgcn_ballot_w64(!(dmax < P.lim)) != 0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!(pfast & (d[u] < P.lim))) {
          const int gr = grid + u;
          if (nrec < kGatherSlots) {
            recs[nrec * nth + lane] = (unsigned)gr;
            F[u] = make_float4(0.f, 0.f, 0.f, 0.f);
          } else {
            const int ex = cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half,
                                           P.h, P.lg);
            F[u] = cell_load(force + (long long)gr * P.cells, ex);
          }
          ++nrec;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (WORLD) {
        tx += F[u].x;
        ty += F[u].y;
        tz += F[u].z;
      } else {
        accumulate_rotated(rot + 12 * (grid + u), F[u], tx, ty, tz);
      }
    }
  }
  for (; grid < P.k; ++grid) {
    const int cell =
        cell_index(c0, c1, c2, pfast, frot + 12 * grid, rot + 12 * grid, off + 4 * grid, P);
    const float4 F = cell_load(force + (long long)grid * P.cells, cell);
    if (WORLD) {
      tx += F.x;
      ty += F.y;
      tz += F.z;
    } else {
      accumulate_rotated(rot + 12 * grid, F, tx, ty, tz);
    }
  }
  // the deferred pairs, in the order they were met
  const int mine = min(nrec, kGatherSlots);
  for (int r = 0; r < mine; ++r) {
    const int gr = (int)recs[r * nth + lane];
    const int ex =
        cell_exact_call(c0, c1, c2, rot + 12 * gr, off + 4 * gr, P.W, P.half, P.h, P.lg);
    const float4 F = cell_load(force + (long long)gr * P.cells, ex);
    if (WORLD) {
      tx += F.x;
      ty += F.y;
      tz += F.z;
    } else {
      accumulate_rotated(rot + 12 * gr, F, tx, ty, tz);
    }
  }
}

}  // namespace mog
