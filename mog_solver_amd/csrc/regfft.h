// In-register FFTs of one length-G line per thread (G = 16, 32, 64), used
// by the spectral solve's slab passes.  Radix-2 decimation in time: the
// input permutation and every twiddle index are compile-time constants once
// the loops are unrolled, so the permutation is register renaming, the
// trivial twiddles (1 and -+i) cost no multiply, and the others are
// immediate constants.  Twiddles are the FP64 values of (cos, sin)(2 pi j /
// 64) rounded to float32, the same values as the tables the other paths
// build on the host.  Forward: X_k = sum_n x_n e^{-2 pi i n k / G}; inverse:
// e^{+...}, unnormalised.
#pragma once

#include <hip/hip_runtime.h>

namespace mog {

// (cos, sin)(2 pi j / 64)
__device__ constexpr float kTw64[64][2] = {
    {0x1.0000000000000p+0f, 0x0.0p+0f},
    {0x1.fd88da0000000p-1f, 0x1.917a6c0000000p-4f},
    {0x1.f6297c0000000p-1f, 0x1.8f8b840000000p-3f},
    {0x1.e9f4160000000p-1f, 0x1.2940620000000p-2f},
    {0x1.d906bc0000000p-1f, 0x1.87de2a0000000p-2f},
    {0x1.c38b300000000p-1f, 0x1.e2b5d40000000p-2f},
    {0x1.a9b6620000000p-1f, 0x1.1c73b40000000p-1f},
    {0x1.8bc8060000000p-1f, 0x1.44cf320000000p-1f},
    {0x1.6a09e60000000p-1f, 0x1.6a09e60000000p-1f},
    {0x1.44cf320000000p-1f, 0x1.8bc8060000000p-1f},
    {0x1.1c73b40000000p-1f, 0x1.a9b6620000000p-1f},
    {0x1.e2b5d40000000p-2f, 0x1.c38b300000000p-1f},
    {0x1.87de2a0000000p-2f, 0x1.d906bc0000000p-1f},
    {0x1.2940620000000p-2f, 0x1.e9f4160000000p-1f},
    {0x1.8f8b840000000p-3f, 0x1.f6297c0000000p-1f},
    {0x1.917a6c0000000p-4f, 0x1.fd88da0000000p-1f},
    {0x1.1a62640000000p-54f, 0x1.0000000000000p+0f},
    {-0x1.917a6c0000000p-4f, 0x1.fd88da0000000p-1f},
    {-0x1.8f8b840000000p-3f, 0x1.f6297c0000000p-1f},
    {-0x1.2940620000000p-2f, 0x1.e9f4160000000p-1f},
    {-0x1.87de2a0000000p-2f, 0x1.d906bc0000000p-1f},
    {-0x1.e2b5d40000000p-2f, 0x1.c38b300000000p-1f},
    {-0x1.1c73b40000000p-1f, 0x1.a9b6620000000p-1f},
    {-0x1.44cf320000000p-1f, 0x1.8bc8060000000p-1f},
    {-0x1.6a09e60000000p-1f, 0x1.6a09e60000000p-1f},
    {-0x1.8bc8060000000p-1f, 0x1.44cf320000000p-1f},
    {-0x1.a9b6620000000p-1f, 0x1.1c73b40000000p-1f},
    {-0x1.c38b300000000p-1f, 0x1.e2b5d40000000p-2f},
    {-0x1.d906bc0000000p-1f, 0x1.87de2a0000000p-2f},
    {-0x1.e9f4160000000p-1f, 0x1.2940620000000p-2f},
    {-0x1.f6297c0000000p-1f, 0x1.8f8b840000000p-3f},
    {-0x1.fd88da0000000p-1f, 0x1.917a6c0000000p-4f},
    {-0x1.0000000000000p+0f, 0x1.1a62640000000p-53f},
    {-0x1.fd88da0000000p-1f, -0x1.917a6c0000000p-4f},
    {-0x1.f6297c0000000p-1f, -0x1.8f8b840000000p-3f},
    {-0x1.e9f4160000000p-1f, -0x1.2940620000000p-2f},
    {-0x1.d906bc0000000p-1f, -0x1.87de2a0000000p-2f},
    {-0x1.c38b300000000p-1f, -0x1.e2b5d40000000p-2f},
    {-0x1.a9b6620000000p-1f, -0x1.1c73b40000000p-1f},
    {-0x1.8bc8060000000p-1f, -0x1.44cf320000000p-1f},
    {-0x1.6a09e60000000p-1f, -0x1.6a09e60000000p-1f},
    {-0x1.44cf320000000p-1f, -0x1.8bc8060000000p-1f},
    {-0x1.1c73b40000000p-1f, -0x1.a9b6620000000p-1f},
    {-0x1.e2b5d40000000p-2f, -0x1.c38b300000000p-1f},
    {-0x1.87de2a0000000p-2f, -0x1.d906bc0000000p-1f},
    {-0x1.2940620000000p-2f, -0x1.e9f4160000000p-1f},
    {-0x1.8f8b840000000p-3f, -0x1.f6297c0000000p-1f},
    {-0x1.917a6c0000000p-4f, -0x1.fd88da0000000p-1f},
    {-0x1.a793940000000p-53f, -0x1.0000000000000p+0f},
    {0x1.917a6c0000000p-4f, -0x1.fd88da0000000p-1f},
    {0x1.8f8b840000000p-3f, -0x1.f6297c0000000p-1f},
    {0x1.2940620000000p-2f, -0x1.e9f4160000000p-1f},
    {0x1.87de2a0000000p-2f, -0x1.d906bc0000000p-1f},
    {0x1.e2b5d40000000p-2f, -0x1.c38b300000000p-1f},
    {0x1.1c73b40000000p-1f, -0x1.a9b6620000000p-1f},
    {0x1.44cf320000000p-1f, -0x1.8bc8060000000p-1f},
    {0x1.6a09e60000000p-1f, -0x1.6a09e60000000p-1f},
    {0x1.8bc8060000000p-1f, -0x1.44cf320000000p-1f},
    {0x1.a9b6620000000p-1f, -0x1.1c73b40000000p-1f},
    {0x1.c38b300000000p-1f, -0x1.e2b5d40000000p-2f},
    {0x1.d906bc0000000p-1f, -0x1.87de2a0000000p-2f},
    {0x1.e9f4160000000p-1f, -0x1.2940620000000p-2f},
    {0x1.f6297c0000000p-1f, -0x1.8f8b840000000p-3f},
    {0x1.fd88da0000000p-1f, -0x1.917a6c0000000p-4f},
};

__host__ __device__ constexpr int bitrev(int i, int lg) {
  int r = 0;
  for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
  return r;
}

template <int G>
__host__ __device__ constexpr int log2_of() {
  return G == 16 ? 4 : G == 32 ? 5 : G == 64 ? 6 : G == 8 ? 3 : G == 4 ? 2 : 1;
}

template <int G, bool INV>
__device__ __forceinline__ void reg_fft(float2 (&x)[G]) {
  constexpr int LG = log2_of<G>();
  static_assert((1 << LG) == G && G <= 64, "G must be a power of two <= 64");
  float2 a[G];
#pragma unroll
  for (int i = 0; i < G; ++i) a[i] = x[bitrev(i, LG)];
#pragma unroll
  for (int s = 1; s <= LG; ++s) {
    const int len = 1 << s, h = len >> 1, step = 64 >> s;
#pragma unroll
    for (int i = 0; i < G; i += len) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        float2 t = a[i + j + h];
        if (j == 0) {
        } else if (4 * j == len) {  // w = -i (forward) or +i (inverse)
          t = INV ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);
        } else {
          const float c = kTw64[j * step][0];
          const float sn = INV ? kTw64[j * step][1] : -kTw64[j * step][1];
          t = make_float2(fmaf(t.x, c, -(t.y * sn)), fmaf(t.x, sn, t.y * c));
        }
        const float2 u = a[i + j];
        a[i + j] = make_float2(u.x + t.x, u.y + t.y);
        a[i + j + h] = make_float2(u.x - t.x, u.y - t.y);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < G; ++i) x[i] = a[i];
}

}  // namespace mog
