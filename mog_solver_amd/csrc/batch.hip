// Batched ensemble of independent MOG simulations: the gather+update as one
// persistent launch over all simulations (the batch's deposit lives beside
// the single-simulation deposit in deposit.hip, its solve beside the
// single-simulation solve in spectral.hip).
//
// Gather+update: a grid-stride loop over blocks of particles that never
// straddle two simulations; a block finds its simulation by binary search
// over the block offsets and then runs the single-simulation world-frame
// gather (same summation order, same integrator), so the new state is
// bit-identical to stepping each simulation alone.
#include "gather_sum.h"

namespace mog {

// Blocks of kBatchGatherBlock particles, each inside one simulation, so the
// simulation (and with it every transform row) is uniform across the
// workgroup and comes through scalar loads; each block runs the
// single-simulation gather's grid loop (gather_sum: shifted-field cells,
// addends staged in LDS, deferral of undecided pairs), so its summation
// order and results are the single-simulation step's.
template <bool LDS3>
__global__ __launch_bounds__(kBatchGatherBlock) __attribute__((amdgpu_waves_per_eu(8)))
void batch_gather_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, long long blocks,
    const BatchSim* __restrict__ sims, int nsims, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    const float4* __restrict__ force, const int* __restrict__ order, float4* __restrict__ ps,
    float4* __restrict__ vs) {
  __shared__ unsigned recs[kGatherSlots * kBatchGatherBlock];
  extern __shared__ float adds[];  // LDS3: the block's simulation's addends (3 per grid)
  for (long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const BatchSim& b = sims[sim_of_block(sims, nsims, blk)];
    const Params P = b.P;
    const float* FR = frot + 12 * b.grid_off;
    if (LDS3) {
      __syncthreads();  // the previous block's reads are done
      for (int i = threadIdx.x; i < 3 * P.k; i += blockDim.x) adds[i] = FR[12 * (i / 3) + 4 * (i % 3) + 3];
      __syncthreads();
    }
    const long long j = (blk - b.blk_off) * kBatchGatherBlock + threadIdx.x;
    if (j >= b.n) continue;
    // the order permutes particles inside the simulation's own rows; with
    // ps, vs the state in visiting order is read and written at the slot
    const long long slot = b.p_off + j;
    const long long i = order ? (long long)order[slot] : slot;
    const float* R = rot + 12 * b.grid_off;
    const float* O = off + 4 * b.grid_off;
    const float4* F = force + b.cell_off;
    float4 p = ps ? ps[slot] : pos[i];
    const float c0 = p.x - P.half, c1 = p.y - P.half, c2 = p.z - P.half;
    const int pfast = particle_fast(c0, c1, c2, P.CB);
    float tx = 0.f, ty = 0.f, tz = 0.f;
    if (P.shift_ok & kShiftF4)
      gather_sum<true, LDS3, true>(c0, c1, c2, pfast, P, R, O, FR, F, adds, recs,
                                   kBatchGatherBlock, threadIdx.x, tx, ty, tz);
    else
      gather_sum<true, LDS3, false>(c0, c1, c2, pfast, P, R, O, FR, F, adds, recs,
                                    kBatchGatherBlock, threadIdx.x, tx, ty, tz);
    float4 v = vs ? vs[slot] : vel[i];
    integrate(p, v, tx / P.fk, ty / P.fk, tz / P.fk, P);
    pos[i] = p;
    if (ps) {
      ps[slot] = p;
      vs[slot] = v;
    } else {
      vel[i] = v;
    }
  }
}

hipError_t launch_batch_gather(float4* pos, float4* vel, long long blocks, const BatchSim* sims,
                               int nsims, int kmax, const float* rot, const float* off,
                               const float* frot, const float4* force, const int* order,
                               hipStream_t s, float4* ps, float4* vs) {
  if (blocks <= 0) return hipSuccess;
  if ((ps || vs) && !(ps && vs && order)) return hipErrorInvalidValue;
  const long long grid = blocks < 16384 ? blocks : 16384;
  const size_t lds = (size_t)kmax * 3 * sizeof(float);
  if (lds <= 16 * 1024) {  // k <= 1365: a few KB per workgroup
    hipLaunchKernelGGL(batch_gather_kernel<true>, dim3((unsigned)grid), dim3(kBatchGatherBlock),
                       lds, s, pos, vel, blocks, sims, nsims, rot, off, frot, force, order, ps,
                       vs);
  } else {
    hipLaunchKernelGGL(batch_gather_kernel<false>, dim3((unsigned)grid), dim3(kBatchGatherBlock),
                       0, s, pos, vel, blocks, sims, nsims, rot, off, frot, force, order, ps,
                       vs);
  }
  return hipGetLastError();
}

}  // namespace mog
