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

__global__ __launch_bounds__(1024) void batch_deposit_kernel(
    const float4* __restrict__ pos, const BatchSim* __restrict__ sims,
    const DepositTask* __restrict__ tasks, int ntasks, const float* __restrict__ rot,
    const float* __restrict__ off, const float* __restrict__ frot,
    unsigned long long* __restrict__ grids, unsigned long long* __restrict__ counter) {
  extern __shared__ unsigned int hist[];
  __shared__ int cur;
  for (;;) {
    if (threadIdx.x == 0) cur = (int)atomicAdd(counter, 1ull);
    __syncthreads();
    const int ti = cur;
    if (ti >= ntasks) break;  // uniform across the workgroup
    const DepositTask task = tasks[ti];
    const BatchSim& b = sims[task.sim];
    const Params P = b.P;
    const int cells = P.cells;
    const int total = task.ng * cells;
    for (int i = threadIdx.x; i < total; i += blockDim.x) hist[i] = 0u;
    __syncthreads();
    const long long p0 = b.p_off + task.p0, p1 = b.p_off + task.p1;
    const unsigned long long lds_limit = 0xFFFFFFFFull / (unsigned long long)(task.p1 - task.p0);
    const int gbase = b.grid_off + task.grid0;
    unsigned long long* dst = grids + b.cell_off + (long long)task.grid0 * cells;
    for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
      const float4 q = pos[p];
      const unsigned long long milli = milli_of(q.w);
      if (milli == 0ull) continue;
      const float c0 = q.x - P.half, c1 = q.y - P.half, c2 = q.z - P.half;
      const int pfast = particle_fast(c0, c1, c2, P.CB);
      for (int gi = 0; gi < task.ng; ++gi) {
        const int gr = gbase + gi;
        const int cell = cell_index(c0, c1, c2, pfast, frot + 12 * gr, rot + 12 * gr, off + 4 * gr, P);
        if (milli <= lds_limit)
          atomicAdd(&hist[gi * cells + cell], (unsigned int)milli);
        else
          atomicAdd(&dst[(long long)gi * cells + cell], milli);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const unsigned int v = hist[i];
      if (v) atomicAdd(&dst[i], (unsigned long long)v);
    }
    __syncthreads();  // LDS and `cur` are reused by the next task
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
    for (int gr = 0; gr < P.k; ++gr) {
      const int cell = cell_index(c0, c1, c2, pfast, FR + 12 * gr, R + 12 * gr, O + 4 * gr, P);
      const float4 f = F[(long long)gr * P.cells + cell];
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
