// Workgroup-to-tile order of the step's launches, shared by the device
// kernels and a host-side check (plain C++, no HIP types).
//
// Workgroup L of a launch runs on XCD L % 8: dispatch deals workgroups
// round-robin over the eight XCDs of an MI355X, each with its own L2.
// xcd_tile maps L to a tile number so that every XCD walks one contiguous
// range of the T tiles, in order of L: tiles that are neighbours in tile
// order, and share inputs, then meet in the same L2. XCD x owns
// q + (x < r) tiles (q = T / 8, r = T % 8), starting at x q + min(x, r);
// the map is a bijection on [0, T) for every T.
#pragma once

#if defined(__HIPCC__)
#define MOG_TILING_HD __host__ __device__ __forceinline__
#else
#define MOG_TILING_HD inline
#endif

namespace mog {

MOG_TILING_HD unsigned int xcd_tile(unsigned int L, unsigned int T) {
  const unsigned int x = L & 7u, s = L >> 3, q = T >> 3, r = T & 7u;
  return x * q + (x < r ? x : r) + s;
}

}  // namespace mog
