// Force gather + integrator of the MOG step (replaces shaders/update.glsl).
//
// One thread per particle reads its position and velocity once, loops over
// the k grids (the per-grid fast rows are wave-uniform, so they come through
// the scalar cache), recomputes the NGP cell exactly as the deposit does
// (same cell for the same position), fetches the force of that cell in the
// world frame (M F: transpose(rot) * gridForce of update.glsl:77), averages
// over k and applies kick, damping, drift and the periodic wrap
// (update.glsl:81-89).
#include "gather_sum.h"

namespace mog {

// mode 0: update pos/vel in place; mode 1: write avgForce to avg_out.
// `order` (optional) is the visiting order: thread t handles particle
// order[t], so that the lanes of a wave are neighbours in space.  ps, vs
// (optional, mode 0 with an order): the state in visiting order, read and
// written at slot t; the new position also goes to pos[order[t]], vel is
// not touched.  LDS3: the fast rows' addends staged in LDS (when k * 12
// bytes fit the launch's LDS budget); the lanes' records of deferred pairs
// follow them.
template <bool WORLD, bool LDS3, bool SH>
__global__ __launch_bounds__(256) void gather_update_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, long long n,
    const float* __restrict__ rot, const float* __restrict__ off, const float* __restrict__ frot,
    const float4* __restrict__ force, Params P, float4* __restrict__ avg_out, int mode,
    const int* __restrict__ order, long long* __restrict__ zero, long long zero_count,
    float4* __restrict__ ps, float4* __restrict__ vs, float4* __restrict__ cpos_out,
    unsigned int lds_limit32) {
  extern __shared__ float adds[];  // LDS3: adds[3 g + m] = frot[12 g + 4 m + 3]; then the records
  unsigned* const recs = reinterpret_cast<unsigned*>(adds + (LDS3 ? 3 * P.k : 0));
  // side job: clear `zero` (the density grids the next step deposits into)
  for (long long z = (long long)blockIdx.x * blockDim.x + threadIdx.x; z < zero_count;
       z += (long long)gridDim.x * blockDim.x)
    zero[z] = 0;
  if (LDS3) {
    for (int i = threadIdx.x; i < 3 * P.k; i += blockDim.x) adds[i] = frot[12 * (i / 3) + 4 * (i % 3) + 3];
    __syncthreads();
  }
  // neighbouring slots of the visiting order on one XCD: they share grid cells
  const long long t = (long long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long i = order ? (long long)order[t] : t;
  float4 p = ps ? ps[t] : pos[i];
  const float c0 = p.x - P.half;
  const float c1 = p.y - P.half;
  const float c2 = p.z - P.half;
  const int pfast = particle_fast(c0, c1, c2, P.CB);
  float tx = 0.f, ty = 0.f, tz = 0.f;
  gather_sum<WORLD, LDS3, SH>(c0, c1, c2, pfast, P, rot, off, frot, force, adds, recs, blockDim.x,
                              threadIdx.x, tx, ty, tz);
  const float ax = tx / P.fk, ay = ty / P.fk, az = tz / P.fk;
  if (mode == 1) {
    avg_out[i] = make_float4(ax, ay, az, 0.0f);
    return;
  }
  float4 v = vs ? vs[t] : vel[i];
  integrate(p, v, ax, ay, az, P);
  if (cpos_out)
    cpos_out[i] = prepared_row(p, P, lds_limit32);
  else
    pos[i] = p;
  if (ps) {
    ps[t] = p;
    vs[t] = v;
  } else {
    vel[i] = v;
  }
}

__global__ __launch_bounds__(256) void state_import_kernel(const float4* __restrict__ pos,
                                                           const float4* __restrict__ vel,
                                                           const int* __restrict__ order,
                                                           long long n, float4* __restrict__ ps,
                                                           float4* __restrict__ vs) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long i = order[t];
  ps[t] = pos[i];
  vs[t] = vel[i];
}

__global__ __launch_bounds__(256) void state_export_kernel(const float4* __restrict__ vs,
                                                           const int* __restrict__ order,
                                                           long long n, float4* __restrict__ vel,
                                                           const float4* __restrict__ ps,
                                                           float4* __restrict__ pos) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long i = order[t];
  vel[i] = vs[t];
  if (ps) pos[i] = ps[t];
}

hipError_t launch_state_import(const float4* pos, const float4* vel, const int* order, long long n,
                               float4* ps, float4* vs, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(state_import_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                     vel, order, n, ps, vs);
  return hipGetLastError();
}

hipError_t launch_state_export(const float4* vs, const int* order, long long n, float4* vel,
                               hipStream_t s, const float4* ps, float4* pos) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(state_export_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vs,
                     order, n, vel, ps, pos);
  return hipGetLastError();
}

// Gather + update from the cell cache of the step's sorted deposit: slot t
// reads its k cells with coalesced uint16 loads (no cell arithmetic), the
// world-frame forces of those cells, then updates particle order[t].  Pairs
// the fast path could not decide (bit 15 of the cached cell) are summed as
// gather_update_kernel sums them: the first kGatherSlots of them in the
// grids below k - k % 4 after all others, the rest in grid order, so both
// gathers give the same bits.
__global__ __launch_bounds__(256) void gather_cached_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, long long n, const int* __restrict__ order,
    const unsigned short* __restrict__ cells, const float4* __restrict__ force, Params P) {
  constexpr int U = 8;
  __shared__ unsigned recs[kGatherSlots * 256];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int k4 = P.k & ~3;  // the grids gather_update_kernel takes in groups of 4
  float tx = 0.f, ty = 0.f, tz = 0.f;
  int nrec = 0;
  auto add = [&](int grid, unsigned c) {
    if ((c & 0x8000u) && grid < k4) {
      if (nrec < kGatherSlots) {
        recs[nrec * blockDim.x + threadIdx.x] = (unsigned)grid;
        ++nrec;
        return;
      }
      ++nrec;
    }
    const float4 F = force[(long long)grid * P.cells + (c & 0x7FFFu)];
    tx += F.x;
    ty += F.y;
    tz += F.z;
  };
  int grid = 0;
  for (; grid + U <= P.k; grid += U) {
    unsigned c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = cells[cache_index(t, grid + u, P.k)];
    if (__ballot(((c[0] | c[1] | c[2] | c[3] | c[4] | c[5] | c[6] | c[7]) & 0x8000u) != 0)) {
#pragma unroll
      for (int u = 0; u < U; ++u) add(grid + u, c[u]);
    } else {
      float4 F[U];
#pragma unroll
      for (int u = 0; u < U; ++u) F[u] = force[(long long)(grid + u) * P.cells + c[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        tx += F[u].x;
        ty += F[u].y;
        tz += F[u].z;
      }
    }
  }
  for (; grid < P.k; ++grid) add(grid, cells[cache_index(t, grid, P.k)]);
  const int mine = min(nrec, kGatherSlots);
  for (int r = 0; r < mine; ++r) {
    const int gr = (int)recs[r * blockDim.x + threadIdx.x];
    const float4 F = force[(long long)gr * P.cells + (cells[cache_index(t, gr, P.k)] & 0x7FFFu)];
    tx += F.x;
    ty += F.y;
    tz += F.z;
  }
  const long long i = order[t];
  float4 p = pos[i];
  float4 v = vel[i];
  integrate(p, v, tx / P.fk, ty / P.fk, tz / P.fk, P);
  pos[i] = p;
  vel[i] = v;
}

hipError_t launch_gather_cached(const Params& P, float4* pos, float4* vel, long long n,
                                const int* order, const unsigned short* cells,
                                const float4* force, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_cached_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos,
                     vel, n, order, cells, force, P);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void update_kernel(float4* __restrict__ pos,
                                                     float4* __restrict__ vel,
                                                     const float4* __restrict__ avg, long long n,
                                                     Params P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  float4 v = vel[i];
  const float4 a = avg[i];
  integrate(p, v, a.x, a.y, a.z, P);
  pos[i] = p;
  vel[i] = v;
}

hipError_t launch_gather_update(const Params& P, float4* pos, float4* vel, long long n,
                                const float* rot, const float* off, const float* frot,
                                const float4* force, int world, float4* avg_out, int mode,
                                const int* order, hipStream_t s, long long* zero,
                                long long zero_count, float4* ps, float4* vs,
                                float4* cpos_out, unsigned int lds_limit32) {
  if (n <= 0) return zero_count > 0 ? launch_fill_zero(zero, zero_count, s) : hipSuccess;
  if ((ps || vs) && !(ps && vs && order && mode == 0)) return hipErrorInvalidValue;
  if (cpos_out && !ps) return hipErrorInvalidValue;
  const long long blocks = (n + 255) / 256;
  const size_t lds = (size_t)P.k * 3 * sizeof(float);
  const bool lds3 = lds <= 16 * 1024;  // k <= 1365: a few KB per workgroup
  const size_t lds_recs = (size_t)kGatherSlots * 256 * sizeof(unsigned);
  const bool sh = P.shift_ok & kShiftF4;
  const dim3 grid((unsigned)blocks);
#define MOG_GATHER(W3, L3, S3)                                                             \
  hipLaunchKernelGGL((gather_update_kernel<W3, L3, S3>), grid, dim3(256),                 \
                     (L3 ? lds : 0) + lds_recs, s, pos, vel, n, rot, off, frot, force, P, \
                     avg_out, mode, order, zero, zero_count, ps, vs, cpos_out, lds_limit32)
  if (world && lds3 && sh) MOG_GATHER(true, true, true);
  else if (world && lds3) MOG_GATHER(true, true, false);
  else if (world && sh) MOG_GATHER(true, false, true);
  else if (world) MOG_GATHER(true, false, false);
  else if (lds3 && sh) MOG_GATHER(false, true, true);
  else if (lds3) MOG_GATHER(false, true, false);
  else if (sh) MOG_GATHER(false, false, true);
  else MOG_GATHER(false, false, false);
#undef MOG_GATHER
  return hipGetLastError();
}

hipError_t launch_update(const Params& P, float4* pos, float4* vel, const float4* avg,
                         long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(update_kernel, dim3((unsigned)blocks), dim3(256), 0, s, pos, vel, avg, n, P);
  return hipGetLastError();
}

}  // namespace mog
