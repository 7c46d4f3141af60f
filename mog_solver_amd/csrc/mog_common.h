// Shared device helpers and launch parameters of the MOG force step.
//
// The whole library is compiled with -ffp-contract=off: the cell-index
// arithmetic must round every product and sum separately (same order as the
// oracle's float32 emulation) so that density grids are bit-exact.  Code
// that wants fused multiply-adds writes fmaf() explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mog {

// Scalar parameters, float32 exactly as the reference shaders receive them
// (main.py:604-605, 662-667): W = float(world_size), half = W * 0.5f,
// h = float(world_size / grid_size), dtp = float(dt * timescale).
struct Params {
  int k;       // grids
  int g;       // cells per axis (power of two)
  int lg;      // log2(g)
  int cells;   // g^3
  float W, half, h;
  float dtp, damping;
  float fk;    // float(k), the divisor of update.glsl:81
};

// GLSL mod(x, y) = x - y * floor(x / y), IEEE division, no contraction.
__device__ __forceinline__ float glsl_mod(float x, float y) {
  float t = floorf(x / y);
  float p = y * t;
  return x - p;
}

// One axis of the NGP cell (mass.glsl:46-50): int(mod(mod(l,W)+W,W)/h) % g.
// int() truncates toward zero; the value is >= 0 for finite l, so % g is
// & (g-1).  Non-finite l lands in an in-bounds cell.
__device__ __forceinline__ int axis_cell(float l, const Params& P) {
  l = glsl_mod(l, P.W);
  l = l + P.W;
  l = glsl_mod(l, P.W);
  float q = l / P.h;
  int i = (int)q;
  return i & (P.g - 1);
}

// Flat cell ((z*g)+y)*g+x of the centred position (c0,c1,c2) in the grid
// whose packed rotation row is R (12 floats, R[4j+m] = M[j][m]) and offset
// O (4 floats).  Local position r = mat3(col0,col1,col2) * c, i.e.
// r_m = (R[m]*c0 + R[4+m]*c1) + R[8+m]*c2 (mass.glsl:20-25, 42-44).
__device__ __forceinline__ int cell_of(float c0, float c1, float c2, const float* __restrict__ R,
                                       const float* __restrict__ O, const Params& P) {
  float r0 = R[0] * c0;
  r0 = r0 + R[4] * c1;
  r0 = r0 + R[8] * c2;
  float r1 = R[1] * c0;
  r1 = r1 + R[5] * c1;
  r1 = r1 + R[9] * c2;
  float r2 = R[2] * c0;
  r2 = r2 + R[6] * c1;
  r2 = r2 + R[10] * c2;
  float l0 = (r0 + P.half) - O[0];
  float l1 = (r1 + P.half) - O[1];
  float l2 = (r2 + P.half) - O[2];
  int x = axis_cell(l0, P);
  int y = axis_cell(l1, P);
  int z = axis_cell(l2, P);
  return (((z << P.lg) | y) << P.lg) | x;
}

// uint(mass * 1000.0) of mass.glsl:54 as a 64-bit integer; negative and NaN
// masses deposit 0.
__device__ __forceinline__ unsigned long long milli_of(float m) {
  float mm = m * 1000.0f;
  return mm > 0.0f ? (unsigned long long)mm : 0ull;
}

}  // namespace mog

// Host-side launchers (defined in the kernel translation units).
namespace mog {

struct DepositTune {
  int threads;        // threads per workgroup
  int grids_per_wg;   // grids privatised in LDS per workgroup
  int wg_target;      // aim for about this many workgroups
};

hipError_t launch_fill_zero(long long* p, long long count, hipStream_t s);
hipError_t launch_deposit(const Params& P, const float4* pos, long long n, const float* rot,
                          const float* off, unsigned long long* grids, const DepositTune& t,
                          hipStream_t s);
hipError_t launch_solve(const Params& P, const long long* grids, float2* S1, float2* S2,
                        float4* force, const float2* tw, const float* gtab, const float* kd,
                        int need_x, hipStream_t s);
hipError_t launch_gather_update(const Params& P, float4* pos, float4* vel, long long n,
                                const float* rot, const float* off, const float4* force,
                                float4* avg_out, int mode, hipStream_t s);
hipError_t launch_update(const Params& P, float4* pos, float4* vel, const float4* avg,
                         long long n, hipStream_t s);

}  // namespace mog
