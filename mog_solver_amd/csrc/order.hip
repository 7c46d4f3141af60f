// Spatial visiting order for the gather.
//
// The gather's cost is its force fetches: with particles in arbitrary order
// the 64 lanes of a wave read 64 scattered cells of each grid (about 25
// distinct cache lines per wave and grid).  Visiting particles in Morton
// order of their positions makes the lanes of a wave neighbours in space, so
// they share a few cells per grid.  The order is an index array (the caller's
// arrays are not permuted) rebuilt every few steps; results do not depend on
// it, since every particle's gather and update are independent of the others.
//
// Keys: 10 bits per axis of the position in the periodic box, interleaved
// (x in bit 0), sorted with rocPRIM's device radix sort.
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "mog_common.h"

namespace mog {

__device__ __forceinline__ unsigned spread3(unsigned v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ unsigned quant10(float x, float inv_w) {
  float q = x * inv_w * 1024.0f;
  q = fminf(fmaxf(q, 0.0f), 1023.0f);  // NaN -> 0
  return (unsigned)q;
}

__global__ __launch_bounds__(256) void morton_keys_kernel(const float4* __restrict__ pos,
                                                          long long n, float inv_w,
                                                          unsigned* __restrict__ keys,
                                                          int* __restrict__ idx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i];
  keys[i] = spread3(quant10(p.x, inv_w)) | (spread3(quant10(p.y, inv_w)) << 1) |
            (spread3(quant10(p.z, inv_w)) << 2);
  idx[i] = (int)i;
}

hipError_t order_temp_bytes(long long n, size_t* bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs((void*)nullptr, *bytes, (const unsigned*)nullptr,
                                   (unsigned*)nullptr, (const int*)nullptr, (int*)nullptr,
                                   (unsigned)n, 0, 30, (hipStream_t)0);
}

hipError_t build_order(const float4* pos, long long n, float W, unsigned* keys_a,
                       unsigned* keys_b, int* idx_a, int* order, void* tmp, size_t tmp_bytes,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(morton_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, s, pos, n,
                     1.0f / W, keys_a, idx_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = tmp_bytes;
  return rocprim::radix_sort_pairs(tmp, bytes, keys_a, keys_b, idx_a, order, (unsigned)n, 0, 30, s);
}

// The same order sorted between two caller-provided buffer pairs
// (rocPRIM double buffers: the sort ping-pongs between keys_a/keys_b and
// idx_a/idx_b, so its temporary storage holds only its histograms and
// partition state).  *order receives whichever index buffer holds the
// result.
hipError_t order_temp_bytes_db(long long n, size_t* bytes) {
  *bytes = 0;
  rocprim::double_buffer<unsigned> keys(nullptr, nullptr);
  rocprim::double_buffer<int> vals(nullptr, nullptr);
  return rocprim::radix_sort_pairs((void*)nullptr, *bytes, keys, vals, (unsigned)n, 0, 30,
                                   (hipStream_t)0);
}

hipError_t build_order_db(const float4* pos, long long n, float W, unsigned* keys_a,
                          unsigned* keys_b, int* idx_a, int* idx_b, void* tmp, size_t tmp_bytes,
                          const int** order, hipStream_t s) {
  *order = nullptr;
  if (n <= 0) return hipSuccess;
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(morton_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, s, pos, n,
                     1.0f / W, keys_a, idx_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  rocprim::double_buffer<unsigned> keys(keys_a, keys_b);
  rocprim::double_buffer<int> vals(idx_a, idx_b);
  size_t bytes = tmp_bytes;
  e = rocprim::radix_sort_pairs(tmp, bytes, keys, vals, (unsigned)n, 0, 30, s);
  if (e == hipSuccess) *order = vals.current();
  return e;
}

// Batch keys: (sim << 30) | Morton key in the simulation's own box, over
// the gather blocks (simulation uniform per workgroup).
__global__ __launch_bounds__(kBatchGatherBlock) void batch_morton_keys_kernel(
    const float4* __restrict__ pos, long long blocks, const BatchSim* __restrict__ sims,
    int nsims, unsigned long long* __restrict__ keys, int* __restrict__ idx) {
  for (long long blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    const int si = sim_of_block(sims, nsims, blk);
    const BatchSim& b = sims[si];
    const long long j = (blk - b.blk_off) * kBatchGatherBlock + threadIdx.x;
    if (j >= b.n) continue;
    const long long i = b.p_off + j;
    const float4 p = pos[i];
    const float inv_w = 1.0f / b.P.W;
    const unsigned m = spread3(quant10(p.x, inv_w)) | (spread3(quant10(p.y, inv_w)) << 1) |
                       (spread3(quant10(p.z, inv_w)) << 2);
    keys[i] = ((unsigned long long)si << 30) | m;
    idx[i] = (int)i;
  }
}

static int key_bits(int nsims) {
  int b = 0;
  while ((1 << b) < nsims) ++b;
  return 30 + b;
}

hipError_t batch_order_temp_bytes(long long n, int nsims, size_t* bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs((void*)nullptr, *bytes, (const unsigned long long*)nullptr,
                                   (unsigned long long*)nullptr, (const int*)nullptr, (int*)nullptr,
                                   (unsigned)n, 0, key_bits(nsims), (hipStream_t)0);
}

hipError_t build_batch_order(const float4* pos, long long n, long long blocks,
                             const BatchSim* sims, int nsims, unsigned long long* keys_a,
                             unsigned long long* keys_b, int* idx_a, int* order, void* tmp,
                             size_t tmp_bytes, hipStream_t s) {
  if (n <= 0 || blocks <= 0) return hipSuccess;
  const long long grid = blocks < 16384 ? blocks : 16384;
  hipLaunchKernelGGL(batch_morton_keys_kernel, dim3((unsigned)grid), dim3(kBatchGatherBlock), 0, s,
                     pos, blocks, sims, nsims, keys_a, idx_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = tmp_bytes;
  return rocprim::radix_sort_pairs(tmp, bytes, keys_a, keys_b, idx_a, order, (unsigned)n, 0,
                                   key_bits(nsims), s);
}

}  // namespace mog
