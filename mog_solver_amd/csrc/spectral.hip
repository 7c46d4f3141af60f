// Spectral Poisson solve of the MOG step, FP32 vector path.
//
// Replaces complex.glsl (u32 -> float/1000), the 3-axis radix-2 transforms of
// run_fft_3d_batched (main.py:564-590, fft_batched.glsl), the DC removal of
// fourier.glsl, the Green's multiply and i*k gradients of greens.glsl:25-61,
// the three inverse transforms (main.py:638-646) and pack.glsl.
//
// g = 16 and g = 32 run one workgroup per grid with register FFTs (the
// grid kernels below).  For the other sizes each length-g transform is a
// dense DFT contraction with a twiddle table (FP64-derived, stored FP32) for
// g <= 8, and for g = 64 a radix-8/4/2 Stockham FFT over all lines of the
// slab at once, in LDS (g log2 g instead of g^2 complex multiply-adds per
// line); the 3D transform of a grid is then split into three passes over an
// L2-resident complex scratch buffer:
//   A  per (grid, z) slab: real x-DFT, y-DFT                      -> S1
//   B  per (grid, ky) slab: z-DFT, Green x i*k, inverse z-DFT      -> S1 (y+iz), S2 (x)
//   C  per (grid, z) slab: inverse y-DFT, inverse x-DFT, pack float4
// Two real force components share one complex inverse transform:
// IDFT(A_y + i A_z) = F_y + i F_z, where A_d is the gradient spectrum with
// its Nyquist plane along d removed.  That plane is anti-Hermitian and only
// feeds the imaginary part the reference discards (pack.glsl:23-26), so the
// real parts are unchanged.  The constant factors (1/1000 of complex.glsl,
// -2*pi*G/k^2 of greens.glsl, 1/g^3 of pack.glsl) are folded into one FP64
// table `gtab`.
#include "mog_common.h"
#include "regfft.h"

namespace mog {

// e^{-i theta} * a for the forward sign; tw[j] = (cos, sin)(2*pi*j/g).
__device__ __forceinline__ float2 cmul_conj_tw(float2 a, float2 w) {
  // a * (w.x - i w.y)
  return make_float2(fmaf(a.x, w.x, a.y * w.y), fmaf(a.y, w.x, -(a.x * w.y)));
}
__device__ __forceinline__ float2 cmul_tw(float2 a, float2 w) {
  // a * (w.x + i w.y)
  return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.y, w.x, a.x * w.y));
}

// Grid sizes from which the slab passes use lds_fft instead of dense DFTs.
constexpr int kFftMinG = 16;

// Mixed radix-8/4/2 Stockham FFT (natural order in and out) of `lines`
// independent length-g lines held in LDS, element (line l, position p) at
// p*ps + l*ls.  Stages of radix 8 while at least three factors of two
// remain (but not exactly four: g = 16 runs 4 x 4, g = 32 8 x 4, g = 64
// 8 x 8), then radix 4, then radix 2; with
// Ns the product of the radices of the earlier stages, q = g / R, a stage
// reads one buffer and writes the other:
//   v_r = a[j + r q] * w^r,  w = e^{-+2 pi i (j mod Ns) / (R Ns)},  r < R,
//   (V_0 .. V_{R-1}) = the R-point DFT of v,
//   b[(j / Ns) R Ns + (j mod Ns) + r Ns] = V_r,
// for j < q (forward: e^-, inverse: e^+, unnormalised).  twl holds (cos,
// sin)(2 pi t / g).  Returns the buffer that holds the result.  Called by
// every thread of the workgroup; ends with a barrier.  When the lines are
// the unit-stride direction (ls == 1) consecutive threads take consecutive
// lines, otherwise consecutive positions, so that a wave's LDS accesses
// fall in distinct banks.
__device__ __forceinline__ float2* lds_fft(float2* a, float2* b, int g, int lg, int ps, int ls,
                                           int lines, const float2* twl, bool inverse) {
  const bool line_fast = ls == 1;
  int Ns = 1, rem = lg, lNs = 0;
  while (rem > 0) {
    if (rem >= 3 && rem != 4) {  // radix 8 (g = 32: 8 x 4; g = 64: 8 x 8)
      const int q = g >> 3, lq = lg - 3;
      const int nb = lines * q;
      for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        const int l = line_fast ? t & (lines - 1) : t >> lq;
        const int j = line_fast ? t / lines : t & (q - 1);
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = a[(j + r * q) * ps + l * ls];
        const int k = j & (Ns - 1);
        const int ti = k << (lg - 3 - lNs);  // k * g / (8 Ns)
#pragma unroll
        for (int r = 1; r < 8; ++r) {
          const float2 w = twl[(r * ti) & (g - 1)];
          v[r] = inverse ? cmul_tw(v[r], w) : cmul_conj_tw(v[r], w);
        }
        // 8-point DFT as two 4-point DFTs (even, odd inputs) joined with
        // the eighth roots of unity: V_m = E_m + W^m O_m, V_m+4 = E_m - W^m O_m
        float2 E[4], O[4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float2 x0 = v[h], x1 = v[h + 2], x2 = v[h + 4], x3 = v[h + 6];
          const float2 s02 = make_float2(x0.x + x2.x, x0.y + x2.y);
          const float2 d02 = make_float2(x0.x - x2.x, x0.y - x2.y);
          const float2 s13 = make_float2(x1.x + x3.x, x1.y + x3.y);
          const float2 d13 = make_float2(x1.x - x3.x, x1.y - x3.y);
          const float2 rd13 = inverse ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
          float2* X = h == 0 ? E : O;
          X[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
          X[1] = make_float2(d02.x + rd13.x, d02.y + rd13.y);
          X[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
          X[3] = make_float2(d02.x - rd13.x, d02.y - rd13.y);
        }
        // W^1 = (1 -+ i) c, W^2 = -+i, W^3 = (-1 -+ i) c, c = sqrt(2)/2
        const float c = 0.70710678118654752f;
        const float sg = inverse ? 1.0f : -1.0f;
        float2 t1 = make_float2((O[1].x - sg * O[1].y) * c, (O[1].y + sg * O[1].x) * c);
        float2 t2 = make_float2(-sg * O[2].y, sg * O[2].x);
        float2 t3 = make_float2((-O[3].x - sg * O[3].y) * c, (-O[3].y + sg * O[3].x) * c);
        const float2 T[4] = {O[0], t1, t2, t3};
        const int d = ((j >> lNs) << (lNs + 3)) + k;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          b[(d + m * Ns) * ps + l * ls] = make_float2(E[m].x + T[m].x, E[m].y + T[m].y);
          b[(d + (m + 4) * Ns) * ps + l * ls] = make_float2(E[m].x - T[m].x, E[m].y - T[m].y);
        }
      }
      Ns <<= 3;
      lNs += 3;
      rem -= 3;
    } else if (rem >= 2) {  // radix 4
      const int q = g >> 2, lq = lg - 2;
      const int nb = lines * q;
      for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        const int l = line_fast ? t & (lines - 1) : t >> lq;
        const int j = line_fast ? t / lines : t & (q - 1);
        float2 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = a[(j + r * q) * ps + l * ls];
        const int k = j & (Ns - 1);
        const int ti = k << (lg - 2 - lNs);  // k * g / (4 Ns)
#pragma unroll
        for (int r = 1; r < 4; ++r) {
          const float2 w = twl[(r * ti) & (g - 1)];
          v[r] = inverse ? cmul_tw(v[r], w) : cmul_conj_tw(v[r], w);
        }
        // 4-point DFT; s = -i (forward) or +i (inverse) applied to v1 - v3
        const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
        const float2 rd13 = inverse ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
        const int d = ((j >> lNs) << (lNs + 2)) + k;
        b[d * ps + l * ls] = make_float2(s02.x + s13.x, s02.y + s13.y);
        b[(d + Ns) * ps + l * ls] = make_float2(d02.x + rd13.x, d02.y + rd13.y);
        b[(d + 2 * Ns) * ps + l * ls] = make_float2(s02.x - s13.x, s02.y - s13.y);
        b[(d + 3 * Ns) * ps + l * ls] = make_float2(d02.x - rd13.x, d02.y - rd13.y);
      }
      Ns <<= 2;
      lNs += 2;
      rem -= 2;
    } else {  // radix 2
      const int half = g >> 1;
      const int nb = lines * half;
      for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        const int l = line_fast ? t & (lines - 1) : t >> (lg - 1);
        const int j = line_fast ? t / lines : t & (half - 1);
        const float2 v0 = a[j * ps + l * ls];
        const float2 u = a[(j + half) * ps + l * ls];
        const int k = j & (Ns - 1);
        const float2 w = twl[k << (lg - 1 - lNs)];
        const float2 v1 = inverse ? cmul_tw(u, w) : cmul_conj_tw(u, w);
        const int d = ((j >> lNs) << (lNs + 1)) + k;
        b[d * ps + l * ls] = make_float2(v0.x + v1.x, v0.y + v1.y);
        b[(d + Ns) * ps + l * ls] = make_float2(v0.x - v1.x, v0.y - v1.y);
      }
      Ns <<= 1;
      lNs += 1;
      rem -= 1;
    }
    __syncthreads();
    float2* c = a;
    a = b;
    b = c;
  }
  return a;
}

// Pass A on slab `slab` = (grid, z) of one simulation's arrays.  The pass
// bodies are device functions so that the batched ensemble's persistent
// kernels run exactly the same arithmetic as the single-simulation launch.
__device__ __forceinline__ void spec_fwd_xy_slab(const long long* __restrict__ mass,
                                                 float2* __restrict__ S,
                                                 const float2* __restrict__ tw, int g, int lg,
                                                 long long slab) {
  extern __shared__ float2 sm[];
  float2* twl = sm;
  const int gg = g * g, m = g - 1;
  const long long base = slab * gg;
  if (g >= kFftMinG) {  // [y][x] -> x-FFT -> y-FFT -> S[ky][kx]
    float2* A = sm + g;
    float2* B2 = A + gg;
    for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
    for (int i = threadIdx.x; i < gg; i += blockDim.x)
      A[i] = make_float2((float)mass[base + i], 0.f);
    __syncthreads();
    float2* r = lds_fft(A, B2, g, lg, 1, g, g, twl, false);
    r = lds_fft(r, r == A ? B2 : A, g, lg, g, 1, g, twl, false);
    for (int i = threadIdx.x; i < gg; i += blockDim.x) S[base + i] = r[i];
    return;
  }
  float* re = (float*)(sm + g);          // g*g reals
  float2* B = sm + g + (g * g + 1) / 2;  // g*g complex
  for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
  for (int i = threadIdx.x; i < gg; i += blockDim.x) re[i] = (float)mass[base + i];
  __syncthreads();
  // x: B[y][kx] = sum_x re[y][x] e^{-2 pi i x kx / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, kx = o & m;
    float ar = 0.f, ai = 0.f;
    for (int x = 0; x < g; ++x) {
      const float a = re[y * g + x];
      const float2 w = twl[(x * kx) & m];
      ar = fmaf(a, w.x, ar);
      ai = fmaf(-a, w.y, ai);
    }
    B[o] = make_float2(ar, ai);
  }
  __syncthreads();
  // y: S[ky][kx] = sum_y B[y][kx] e^{-2 pi i y ky / g}
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int ky = o >> lg, kx = o & m;
    float2 acc = make_float2(0.f, 0.f);
    for (int y = 0; y < g; ++y) {
      const float2 v = cmul_conj_tw(B[y * g + kx], twl[(y * ky) & m]);
      acc.x += v.x;
      acc.y += v.y;
    }
    S[base + o] = acc;
  }
}

// Pass A: one workgroup per (grid, z) slab.
__global__ __launch_bounds__(256) void spec_fwd_xy_kernel(const long long* __restrict__ mass,
                                                          float2* __restrict__ S,
                                                          const float2* __restrict__ tw, int g,
                                                          int lg) {
  spec_fwd_xy_slab(mass, S, tw, g, lg, blockIdx.x);
}

// Pass B on slab (grid, ky): columns along z for every kx.
__device__ __forceinline__ void spec_z_green_slab(float2* __restrict__ S, float2* __restrict__ S2,
                                                  const float2* __restrict__ tw,
                                                  const float* __restrict__ gtab,
                                                  const float* __restrict__ kd,
                                                  const float2* __restrict__ conv, int g, int lg,
                                                  int need_x, long long slab) {
  extern __shared__ float2 sm[];
  float2* twl = sm;
  const int gg = g * g, m = g - 1;
  const int grid = (int)(slab >> lg);
  const int ky = (int)(slab & m);
  const long long gbase = (long long)grid * g * gg;
  if (g >= kFftMinG) {
    // [z][kx] -> z-FFT -> Green into ZX[kz][2g] (Z row then X row) -> one
    // inverse FFT over the 2g lines of ZX -> S (Z), S2 (X)
    float2* A = sm + g;      // 2 g*g
    float2* B2 = A + 2 * gg;  // 2 g*g
    for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
    for (int i = threadIdx.x; i < gg; i += blockDim.x) {
      const int z = i >> lg, kx = i & m;
      A[i] = S[gbase + (long long)z * gg + ky * g + kx];
    }
    __syncthreads();
    float2* r = lds_fft(A, B2, g, lg, g, 1, g, twl, false);
    float2* ZX = r == A ? B2 : A;
    const float kyv = kd[ky];
    for (int o = threadIdx.x; o < gg; o += blockDim.x) {
      const int kz = o >> lg, kx = o & m;
      const float2 acc = r[o];
      const int fi = (kz * g + ky) * g + kx;
      float2 zv, xv = make_float2(0.f, 0.f);
      if (conv) {
        const float2 t = conv[fi], u = conv[g * gg + fi];
        zv = make_float2(fmaf(acc.x, t.x, -(acc.y * t.y)), fmaf(acc.x, t.y, acc.y * t.x));
        xv = make_float2(fmaf(acc.x, u.x, -(acc.y * u.y)), fmaf(acc.x, u.y, acc.y * u.x));
      } else {
        const float gf = gtab[fi];
        const float pr = acc.x * gf, pi = acc.y * gf;
        const float kzv = kd[kz];
        zv = make_float2(fmaf(-pr, kzv, -(pi * kyv)), fmaf(pr, kyv, -(pi * kzv)));
        if (need_x) {
          const float kxv = kd[kx];
          xv = make_float2(-(pi * kxv), pr * kxv);
        }
      }
      ZX[kz * 2 * g + kx] = zv;
      ZX[kz * 2 * g + g + kx] = xv;
    }
    __syncthreads();
    // with conv both halves are needed; without, X only when need_x
    const int lines = (conv || need_x) ? 2 * g : g;
    r = lds_fft(ZX, r, g, lg, 2 * g, 1, lines, twl, true);
    for (int o = threadIdx.x; o < gg; o += blockDim.x) {
      const int z = o >> lg, kx = o & m;
      const long long at = gbase + (long long)z * gg + ky * g + kx;
      S[at] = r[z * 2 * g + kx];
      if (need_x) S2[at] = r[z * 2 * g + g + kx];
    }
    return;
  }
  float2* L = sm + g;           // [z][kx]
  float2* Z = L + g * g;        // [kz][kx]  (y + i z spectrum)
  float2* X = Z + g * g;        // [kz][kx]  (x spectrum)
  for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    const int z = i >> lg, kx = i & m;
    L[i] = S[gbase + (long long)z * gg + ky * g + kx];
  }
  __syncthreads();
  const float kyv = kd[ky];
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int kz = o >> lg, kx = o & m;
    float2 acc = make_float2(0.f, 0.f);
    for (int z = 0; z < g; ++z) {
      const float2 v = cmul_conj_tw(L[z * g + kx], twl[(z * kz) & m]);
      acc.x += v.x;
      acc.y += v.y;
    }
    const int fi = (kz * g + ky) * g + kx;
    if (conv) {
      // real-space kernel spectrum: Z = rho * (T_y + i T_z), X = rho * T_x
      const float2 t = conv[fi], u = conv[g * gg + fi];
      Z[o] = make_float2(fmaf(acc.x, t.x, -(acc.y * t.y)), fmaf(acc.x, t.y, acc.y * t.x));
      X[o] = make_float2(fmaf(acc.x, u.x, -(acc.y * u.y)), fmaf(acc.x, u.y, acc.y * u.x));
      continue;
    }
    const float gf = gtab[fi];
    const float pr = acc.x * gf, pi = acc.y * gf;
    // Z = phi * (-kz' + i ky')
    const float kzv = kd[kz];
    Z[o] = make_float2(fmaf(-pr, kzv, -(pi * kyv)), fmaf(pr, kyv, -(pi * kzv)));
    if (need_x) {
      const float kxv = kd[kx];
      X[o] = make_float2(-(pi * kxv), pr * kxv);
    }
  }
  __syncthreads();
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int z = o >> lg, kx = o & m;
    float2 az = make_float2(0.f, 0.f), ax = make_float2(0.f, 0.f);
    for (int kz = 0; kz < g; ++kz) {
      const float2 w = twl[(z * kz) & m];
      const float2 v = cmul_tw(Z[kz * g + kx], w);
      az.x += v.x;
      az.y += v.y;
      if (need_x) {
        const float2 u = cmul_tw(X[kz * g + kx], w);
        ax.x += u.x;
        ax.y += u.y;
      }
    }
    const long long at = gbase + (long long)z * gg + ky * g + kx;
    S[at] = az;
    if (need_x) S2[at] = ax;
  }
}

// Pass B: one workgroup per (grid, ky).
__global__ __launch_bounds__(256) void spec_z_green_kernel(float2* __restrict__ S,
                                                           float2* __restrict__ S2,
                                                           const float2* __restrict__ tw,
                                                           const float* __restrict__ gtab,
                                                           const float* __restrict__ kd,
                                                           const float2* __restrict__ conv,
                                                           int g, int lg, int need_x) {
  spec_z_green_slab(S, S2, tw, gtab, kd, conv, g, lg, need_x, blockIdx.x);
}

// Pass C on slab (grid, z): inverse y, inverse x, pack float4.
// world != 0 stores M F (the grid's force rotated to the world frame,
// world_m = sum_j R[4m+j] F_j) for the step's gather; |F| is unchanged.
__device__ __forceinline__ void spec_inv_xy_slab(const float2* __restrict__ S,
                                                 const float2* __restrict__ S2,
                                                 float4* __restrict__ force,
                                                 const float2* __restrict__ tw, int g, int lg,
                                                 int need_x, const float* __restrict__ rot,
                                                 int world, long long slab) {
  extern __shared__ float2 sm[];
  float2* twl = sm;
  const int gg = g * g, m = g - 1;
  const long long base = slab * gg;
  if (g >= kFftMinG) {
    // ZX[ky][2g] (Z row then X row) -> inverse FFT over ky (2g lines) ->
    // inverse FFT over kx of the 2g rows (row 2y = Z, 2y+1 = X of line y)
    float2* A = sm + g;       // 2 g*g
    float2* B2 = A + 2 * gg;  // 2 g*g
    for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
    for (int i = threadIdx.x; i < gg; i += blockDim.x) {
      const int ky = i >> lg, kx = i & m;
      A[ky * 2 * g + kx] = S[base + i];
      A[ky * 2 * g + g + kx] = need_x ? S2[base + i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int lines = need_x ? 2 * g : g;
    float2* r = lds_fft(A, B2, g, lg, 2 * g, 1, lines, twl, true);
    if (need_x) {
      r = lds_fft(r, r == A ? B2 : A, g, lg, 1, g, 2 * g, twl, true);
    } else {
      r = lds_fft(r, r == A ? B2 : A, g, lg, 1, 2 * g, g, twl, true);
    }
    for (int o = threadIdx.x; o < gg; o += blockDim.x) {
      const int y = o >> lg, x = o & m;
      const float2 tz = r[y * 2 * g + x];
      const float fy = tz.x, fz = tz.y;
      const float fx = need_x ? r[y * 2 * g + g + x].x : fz;
      const float mag = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
      if (world) {
        const float* R = rot + 12 * (slab >> lg);
        const float wx = fmaf(R[0], fx, fmaf(R[1], fy, R[2] * fz));
        const float wy = fmaf(R[4], fx, fmaf(R[5], fy, R[6] * fz));
        const float wz = fmaf(R[8], fx, fmaf(R[9], fy, R[10] * fz));
        force[base + o] = make_float4(wx, wy, wz, mag);
      } else {
        force[base + o] = make_float4(fx, fy, fz, mag);
      }
    }
    return;
  }
  float2* Z = sm + g;       // [ky][kx]
  float2* X = Z + g * g;    // [ky][kx]
  float2* TZ = X + g * g;   // [y][kx]
  float2* TX = TZ + g * g;  // [y][kx]
  for (int i = threadIdx.x; i < g; i += blockDim.x) twl[i] = tw[i];
  for (int i = threadIdx.x; i < gg; i += blockDim.x) {
    Z[i] = S[base + i];
    if (need_x) X[i] = S2[base + i];
  }
  __syncthreads();
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, kx = o & m;
    float2 az = make_float2(0.f, 0.f), ax = make_float2(0.f, 0.f);
    for (int ky = 0; ky < g; ++ky) {
      const float2 w = twl[(y * ky) & m];
      const float2 v = cmul_tw(Z[ky * g + kx], w);
      az.x += v.x;
      az.y += v.y;
      if (need_x) {
        const float2 u = cmul_tw(X[ky * g + kx], w);
        ax.x += u.x;
        ax.y += u.y;
      }
    }
    TZ[o] = az;
    if (need_x) TX[o] = ax;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < gg; o += blockDim.x) {
    const int y = o >> lg, x = o & m;
    float fy = 0.f, fz = 0.f, fx = 0.f;
    for (int kx = 0; kx < g; ++kx) {
      const float2 w = twl[(x * kx) & m];
      const float2 v = cmul_tw(TZ[y * g + kx], w);
      fy += v.x;
      fz += v.y;
      if (need_x) {
        const float2 u = TX[y * g + kx];
        fx = fmaf(u.x, w.x, fmaf(-u.y, w.y, fx));
      }
    }
    if (!need_x) fx = fz;
    const float mag = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
    if (world) {
      const float* R = rot + 12 * (slab >> lg);
      const float wx = fmaf(R[0], fx, fmaf(R[1], fy, R[2] * fz));
      const float wy = fmaf(R[4], fx, fmaf(R[5], fy, R[6] * fz));
      const float wz = fmaf(R[8], fx, fmaf(R[9], fy, R[10] * fz));
      force[base + o] = make_float4(wx, wy, wz, mag);
    } else {
      force[base + o] = make_float4(fx, fy, fz, mag);
    }
  }
}

// Pass C: one workgroup per (grid, z).
__global__ __launch_bounds__(256) void spec_inv_xy_kernel(const float2* __restrict__ S,
                                                          const float2* __restrict__ S2,
                                                          float4* __restrict__ force,
                                                          const float2* __restrict__ tw, int g,
                                                          int lg, int need_x,
                                                          const float* __restrict__ rot,
                                                          int world) {
  spec_inv_xy_slab(S, S2, force, tw, g, lg, need_x, rot, world, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Register-FFT solves for g = 16 and 32: every length-g line is transformed
// by one thread in registers (regfft.h), so an axis costs one LDS read and
// one write per element and no barrier inside it, and the whole solve of a
// grid runs in one workgroup (rf_grid_kernel, g = 16: the full complex
// spectrum in LDS; rh_grid_kernel, g = 32: the half spectrum of the real
// density in LDS, the three force components through a per-workgroup
// scratch slot).
// ---------------------------------------------------------------------------
// Where slab t of a launch lives: its simulation's arrays and tables and
// the slab's index inside that simulation (the grid kernels ask for slab
// grid * g: local / g is the grid).
struct RfSlab {
  const long long* mass;
  float2* S;
  float2* S2;
  float4* force;
  const float* gtab;
  const float* kd;
  const float2* conv;
  const float* rot;
  long long local;
  int need_x;
};

// One simulation: slab t is its slab t.
struct RfSingle {
  const long long* mass;
  float2* S;
  float2* S2;
  float4* force;
  const float* gtab;
  const float* kd;
  const float2* conv;
  const float* rot;
  int need_x;
  __device__ RfSlab operator()(long long t) const {
    return {mass, S, S2, force, gtab, kd, conv, rot, t, need_x};
  }
};

// Two solves of one density in one launch: the first `slabs` slabs (k grids
// x g slabs) are those of `a`, the next `slabs` those of `b` (e.g. two
// Green's functions, two outputs).
struct RfPair {
  RfSingle a, b;
  long long slabs;
  __device__ RfSlab operator()(long long t) const { return t < slabs ? a(t) : b(t - slabs); }
};

// A window of a batch grid-size class: slab t of the class belongs to class
// slot j; its spectral scratch starts at S + soff[j] in the window buffers.
struct RfBatch {
  const BatchSim* sims;
  const int* cls_sim;
  const long long* cls_off;
  int ncls;
  const long long* soff;
  const long long* mass;
  float2* S;
  float2* S2;
  float4* force;
  const float* gtab;
  const float* kd;
  const float2* conv;
  const float* rot;
  __device__ RfSlab operator()(long long t) const;
};

// The spectral multiply of one z line (ky, kx) of a grid's spectrum:
// a[kz] (in: the density spectrum) becomes the packed y + i z force
// spectrum, xs[kz] the x force spectrum; Green's function times i k
// (update.glsl's gradient of the potential) or the convolution solver's two
// precomputed kernels.  gz holds the line's Green's function,
// gtab[(kz G + ky) G + kx] (unused with conv), kdz the wavenumbers kd[kz],
// kyv and kxv kd[ky] and kd[kx], all loaded by the caller.
template <int G>
__device__ __forceinline__ void rf_multiply(const RfSlab& a0, int ky, int kx, float2 (&a)[G],
                                            float2 (&xs)[G], const float (&gz)[G],
                                            const float (&kdz)[G], float kyv, float kxv) {
  constexpr int GG = G * G;
#pragma unroll
  for (int kz = 0; kz < G; ++kz) {
    const int fi = (kz * G + ky) * G + kx;
    const float2 acc = a[kz];
    if (a0.conv) {
      const float2 t = a0.conv[fi], u = a0.conv[G * GG + fi];
      a[kz] = make_float2(fmaf(acc.x, t.x, -(acc.y * t.y)), fmaf(acc.x, t.y, acc.y * t.x));
      xs[kz] = make_float2(fmaf(acc.x, u.x, -(acc.y * u.y)), fmaf(acc.x, u.y, acc.y * u.x));
    } else {
      const float gf = gz[kz];  // a0.gtab[fi], loaded by the caller
      const float pr = acc.x * gf, pi = acc.y * gf;
      const float kzv = kdz[kz];
      a[kz] = make_float2(fmaf(-pr, kzv, -(pi * kyv)), fmaf(pr, kyv, -(pi * kzv)));
      xs[kz] = make_float2(-(pi * kxv), pr * kxv);
    }
  }
}

// The whole solve of one g = 16 grid in one workgroup (512 threads, the two
// force fields of 16^3 complex values in LDS, 69.6 KB): load and convert the
// density, forward FFT along x, y and z, spectral multiply and inverse z
// (pass A and B's work), inverse y and x of both fields and the packed
// float4 store (pass C's).  Each line is transformed by the same register
// FFT, in the same order, as in the three passes, so the result is the
// same; the spectrum never leaves LDS.  Lines: 256 per direction (one per
// thread of the first half); the z pass and the inverse passes give the
// second half the x field.  Block b takes grid g0 + b of `res`.
template <int G, typename R>
__global__ __launch_bounds__(512) void rf_grid_kernel(R res, long long g0, int world) {
  static_assert(G * G == 256, "one line per thread of a half");
  constexpr int GG = G * G, RS = G + 1, PL = G * RS;  // row and plane strides (float2)
  extern __shared__ float2 sm[];
  float2* const A = sm;            // [z][y][x]: the spectrum, then the y + i z field
  float2* const B = sm + G * PL;   // the x field
  const RfSlab a = res((g0 + (long long)blockIdx.x) * G);
  const long long cell0 = a.local * GG;  // the grid's first cell: local = grid * G
  const long long grid = a.local / G;
  const int t = threadIdx.x, half = t >> 8, hi = (t >> 4) & (G - 1), lo = t & (G - 1);
  const bool need_x = a.need_x;
  // the Green's function of this thread's z line (ky = hi, kx = lo), loaded
  // now so that its latency passes under the forward transforms
  float gz[G], kdz[G];
  if (!a.conv) {
#pragma unroll
    for (int kz = 0; kz < G; ++kz) gz[kz] = a.gtab[(kz * G + hi) * G + lo];
  }
#pragma unroll
  for (int kz = 0; kz < G; ++kz) kdz[kz] = a.kd[kz];
  const float kdy = a.kd[hi], kdx = a.kd[lo];
  // density -> complex
  for (int i = t; i < G * GG; i += 512)
    A[(i / GG) * PL + ((i / G) % G) * RS + (i % G)] = make_float2((float)a.mass[cell0 + i], 0.f);
  __syncthreads();
  float2 v[G], xs[G];
  if (half == 0) {  // x lines (z = hi, y = lo)
    float2* row = A + hi * PL + lo * RS;
#pragma unroll
    for (int x = 0; x < G; ++x) v[x] = row[x];
    reg_fft<G, false>(v);
#pragma unroll
    for (int x = 0; x < G; ++x) row[x] = v[x];
  }
  __syncthreads();
  if (half == 0) {  // y lines (z = hi, kx = lo)
    float2* col = A + hi * PL + lo;
#pragma unroll
    for (int y = 0; y < G; ++y) v[y] = col[y * RS];
    reg_fft<G, false>(v);
#pragma unroll
    for (int y = 0; y < G; ++y) col[y * RS] = v[y];
  }
  __syncthreads();
  // z lines (ky = hi, kx = lo): both halves transform the same line, the
  // first keeps the y + i z field, the second the x field
  float2* zl = A + hi * RS + lo;
#pragma unroll
  for (int z = 0; z < G; ++z) v[z] = zl[z * PL];
  __syncthreads();
  if (half == 0 || need_x) {
    reg_fft<G, false>(v);
    rf_multiply<G>(a, hi, lo, v, xs, gz, kdz, kdy, kdx);
    float2* dst = half == 0 ? zl : B + hi * RS + lo;
    if (half == 0) {
      reg_fft<G, true>(v);
#pragma unroll
      for (int z = 0; z < G; ++z) dst[z * PL] = v[z];
    } else {
      reg_fft<G, true>(xs);
#pragma unroll
      for (int z = 0; z < G; ++z) dst[z * PL] = xs[z];
    }
  }
  __syncthreads();
  float2* const F = half == 0 ? A : B;
  const bool busy = half == 0 || need_x;
  if (busy) {  // inverse y (z = hi, kx = lo) of the half's field
    float2* col = F + hi * PL + lo;
#pragma unroll
    for (int ky = 0; ky < G; ++ky) v[ky] = col[ky * RS];
    reg_fft<G, true>(v);
#pragma unroll
    for (int y = 0; y < G; ++y) col[y * RS] = v[y];
  }
  __syncthreads();
  if (busy) {  // inverse x (z = hi, y = lo)
    float2* row = F + hi * PL + lo * RS;
#pragma unroll
    for (int x = 0; x < G; ++x) v[x] = row[x];
    reg_fft<G, true>(v);
#pragma unroll
    for (int x = 0; x < G; ++x) row[x] = v[x];
  }
  __syncthreads();
  const float* Rr = a.rot + 12 * grid;
  float4* out = a.force + cell0;
  for (int i = t; i < G * GG; i += 512) {
    const int at = (i / GG) * PL + ((i / G) % G) * RS + (i % G);
    const float2 tz = A[at];
    const float fy = tz.x, fz = tz.y;
    const float fx = need_x ? B[at].x : fz;
    const float mag = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
    float4 o;
    if (world) {
      o = make_float4(fmaf(Rr[0], fx, fmaf(Rr[1], fy, Rr[2] * fz)),
                      fmaf(Rr[4], fx, fmaf(Rr[5], fy, Rr[6] * fz)),
                      fmaf(Rr[8], fx, fmaf(Rr[9], fy, Rr[10] * fz)), mag);
    } else {
      o = make_float4(fx, fy, fz, mag);
    }
    out[i] = o;
  }
}

// ---------------------------------------------------------------------------
// g = 32: the whole solve of one grid in one workgroup, on the half spectrum.
// The density is real, so its spectrum is fixed by kx = 0 .. 16 (17 x 32 x 32
// complex = 139 KB, which fits the 160 KB of LDS where the full 256 KB
// spectrum does not), and each force component is real, so each comes from
// its own half spectrum.  Per grid:
//   0  density -> float, rows [z][y][x] in LDS (row stride 34 floats)
//   1  real x transforms, two rows per complex FFT:  c = DFT(a + i b),
//      A(k) = (C(k) + conj C(-k)) / 2,  B(k) = (C(k) - conj C(-k)) / 2i,
//      kx = 0 .. 16 written over the rows (17 complex = 34 floats)
//   2  y transforms of the columns kx = 1 .. 15, and of the real columns
//      kx = 0 and 16 packed into one complex FFT and split as in 1
//   3  per z line (ky, kx <= 16): forward z, Green's function (or the
//      convolution multipliers), and for each component d the spectrum
//      i k_d phi (Nyquist entry of k_d zero, as in the other paths),
//      inverse z -> this workgroup's slot of global scratch
//   4  in rounds of 12 z slabs (the slot's slabs copied back into LDS):
//      inverse y (the kx = 0 and 16 columns, Hermitian in ky, packed into
//      one complex FFT whose real and imaginary parts are their two real
//      inverses), inverse x of two rows per complex FFT from their half
//      spectra (Z(k) = X_a(k) + i X_b(k), Z(32 - k) = conj X_a(k) + i
//      conj X_b(k), IDFT Z = x_a + i x_b), then |F| and the packed float4
//      store (world frame: M F).
// The transforms are the unnormalised DFTs of the other paths (the constant
// factors are in gtab), so the forces agree with them to rounding.  Scratch
// per workgroup: three components x 32 slabs x 544 complex (kRhSlot).
// Workgroups loop over grids (grid-stride), one per CU at this LDS size.
// ---------------------------------------------------------------------------
constexpr int kRhThreads = 576;      // 544 z lines in phase 3; 512 tasks elsewhere
constexpr int kRhRow = 34;           // floats per row: 32 reals or 17 complex
// floats per slab (per slab and component in phase 4): 32 rows, padded by 8
// floats so that the slabs or components one wave spans fall in different
// LDS banks (a multiple of 4 floats: the copies of phase 4 store float4)
constexpr int kRhArea = 32 * kRhRow + 8;
constexpr int kRhLines = 32 * 17;    // complex per half-spectrum slab
constexpr long long kRhSlot = 3LL * 32 * kRhLines;  // float2 of scratch per workgroup
constexpr int kRhRound = 12;         // slabs per round of phase 4
constexpr size_t kRhLds = (size_t)kRhRound * 3 * kRhArea * sizeof(float);  // 157,824 B

// The thread index, opaque to the compiler.  rh_grid_kernel takes it afresh
// at the start of each phase, so that the addresses a phase derives from it
// are formed there rather than hoisted out of the grid loop and held in
// registers through every other phase.
__device__ __forceinline__ int phase_index() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// The spectra A, B of two real sequences a, b from C = DFT(a + i b) at k and -k.
__device__ __forceinline__ void split_real_pair(float2 c, float2 cm, float2& A, float2& B) {
  A = make_float2(0.5f * (c.x + cm.x), 0.5f * (c.y - cm.y));
  B = make_float2(0.5f * (c.y + cm.y), 0.5f * (cm.x - c.x));
}

template <typename R>
__global__ __launch_bounds__(kRhThreads) void rh_grid_kernel(R res, long long g0,
                                                             long long ngrids, int world,
                                                             float2* __restrict__ scratch) {
  constexpr int G = 32, H = 17, GG = G * G;
  extern __shared__ float L[];
  float2* const L2 = reinterpret_cast<float2*>(L);
  float2* const slot = scratch + (long long)blockIdx.x * kRhSlot;
  int t;
  for (long long gi = blockIdx.x; gi < ngrids; gi += gridDim.x) {
    const RfSlab a = res((g0 + gi) * G);
    const long long cell0 = a.local * GG;  // local = grid * G
    const long long grid = a.local / G;
    const bool need_x = a.need_x;
    t = phase_index();
    // 0: density -> float rows, in batches of loads that are all in flight
    // before the first conversion waits
    {
      constexpr int NP = G * GG / 2, NL = (NP + kRhThreads - 1) / kRhThreads, NB = 8;
      const longlong2* m2 = reinterpret_cast<const longlong2*>(a.mass + cell0);
#pragma unroll
      for (int r0 = 0; r0 < NL; r0 += NB) {
        longlong2 m[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r) {  // clamped, unconditional loads
          const int i = t + (r0 + r) * kRhThreads;
          if (r0 + r < NL) m[r] = m2[min(i, NP - 1)];
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          const int i = t + (r0 + r) * kRhThreads;
          if (r0 + r < NL && i < NP) {
            const int c = 2 * i, x = c & (G - 1), y = (c >> 5) & (G - 1), z = c >> 10;
            *reinterpret_cast<float2*>(L + z * kRhArea + y * kRhRow + x) =
                make_float2((float)m[r].x, (float)m[r].y);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
    float2 v[G];
    t = phase_index();
    // 1: x transforms, rows p and p + 16 of slab z per thread
    if (t < 512) {
      const int z = t >> 4, p = t & 15;
      float* ra = L + z * kRhArea + p * kRhRow;
      float* rb = ra + 16 * kRhRow;
#pragma unroll
      for (int x = 0; x < G; ++x) v[x] = make_float2(ra[x], rb[x]);
      reg_fft<G, false>(v);
      float2* ha = reinterpret_cast<float2*>(ra);
      float2* hb = reinterpret_cast<float2*>(rb);
#pragma unroll
      for (int k = 0; k <= G / 2; ++k) {
        float2 A, B;
        split_real_pair(v[k], v[(G - k) & (G - 1)], A, B);
        ha[k] = A;
        hb[k] = B;
      }
    }
    __syncthreads();
    t = phase_index();
    // 2: y transforms: threads 0 .. 479 take column q = 1 .. 15 of slab z,
    // threads 480 .. 511 the real columns 0 and 16 of slab z packed into one
    // complex FFT (the two kinds in different waves but one, so that a wave
    // runs one of the two bodies)
    if (t < 480) {
      const int z = t / 15, q = 1 + t - 15 * z;
      float2* col = L2 + z * (kRhArea / 2) + q;
#pragma unroll
      for (int y = 0; y < G; ++y) v[y] = col[y * H];
      reg_fft<G, false>(v);
#pragma unroll
      for (int k = 0; k < G; ++k) col[k * H] = v[k];
    } else if (t < 512) {
      float2* col = L2 + (t - 480) * (kRhArea / 2);
#pragma unroll
      for (int y = 0; y < G; ++y) v[y] = make_float2(col[y * H].x, col[y * H + G / 2].x);
      reg_fft<G, false>(v);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        float2 A, B;
        split_real_pair(v[k], v[(G - k) & (G - 1)], A, B);
        col[k * H] = A;
        col[k * H + G / 2] = B;
      }
    }
    __syncthreads();
    t = phase_index();
    // 3a: forward z transforms of the 544 lines, times Green's function (the
    // convolution solver multiplies per component in 3b), back into LDS.
    // Table and scratch addresses are a wave-uniform base per kz (or z) plus
    // the lane's offset, so that the unrolled loads need no address registers
    // of their own.
    const float2* tb = a.conv ? a.conv + 2 * G * GG : nullptr;  // half-spectrum multipliers
    if (t < kRhLines) {
      const int ky = t / H, kx = t - ky * H;
      int lo = ky * G + kx;
      // opaque to the compiler: the 32 table addresses are formed here, not
      // hoisted out of the grid loop (where they would take 64 registers)
      asm volatile("" : "+v"(lo));
#pragma unroll
      for (int z = 0; z < G; ++z) v[z] = L2[z * (kRhArea / 2) + t];
      reg_fft<G, false>(v);
      if (!tb) {
#pragma unroll
        for (int kz = 0; kz < G; ++kz) {
          const float gf = (a.gtab + kz * GG)[lo];
          v[kz] = make_float2(v[kz].x * gf, v[kz].y * gf);
        }
      }
#pragma unroll
      for (int kz = 0; kz < G; ++kz) L2[kz * (kRhArea / 2) + t] = v[kz];
    }
    __syncthreads();
    t = phase_index();
    // 3b: per component d and line: i k_d times the spectrum, inverse z -> scratch
    if (t < kRhLines) {
      const int ky = t / H, kx = t - ky * H;
      for (int d = need_x ? 0 : 1; d < 3; ++d) {
        if (tb) {
#pragma unroll
          for (int kz = 0; kz < G; ++kz) {
            const float2 u = L2[kz * (kRhArea / 2) + t];
            const float2 m = (tb + (d * G + kz) * kRhLines)[t];
            v[kz] = make_float2(fmaf(u.x, m.x, -(u.y * m.y)), fmaf(u.x, m.y, u.y * m.x));
          }
        } else {
          const float ks = a.kd[d == 0 ? kx : ky];
#pragma unroll
          for (int kz = 0; kz < G; ++kz) {
            const float2 u = L2[kz * (kRhArea / 2) + t];
            const float kv = d == 2 ? a.kd[kz] : ks;
            v[kz] = make_float2(-(u.y * kv), u.x * kv);
          }
        }
        reg_fft<G, true>(v);
#pragma unroll
        for (int z = 0; z < G; ++z) (slot + (d * G + z) * kRhLines)[t] = v[z];
      }
    }
    __syncthreads();
    t = phase_index();
    // 4: inverse y and x, pack
    float4* const out = a.force + cell0;
    const float* Rr = a.rot + 12 * grid;
    for (int z0 = 0; z0 < G; z0 += kRhRound) {
      t = phase_index();  // (per round: addresses not hoisted out of the round loop)
      const int S = min(kRhRound, G - z0);
      constexpr int A4 = kRhLines / 2;  // float4 per slab of scratch
      constexpr int NL = (kRhRound * A4 + kRhThreads - 1) / kRhThreads;  // per component
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // component d of slabs z0 .. z0 + S - 1: a contiguous run of S * A4
        // float4 in the slot, all loads in flight before the first LDS store
        const float4* src = reinterpret_cast<const float4*>(slot + (d * G + z0) * kRhLines);
        float4 c[NL];
#pragma unroll
        for (int r = 0; r < NL; ++r) c[r] = src[min(t + r * kRhThreads, S * A4 - 1)];
#pragma unroll
        for (int r = 0; r < NL; ++r) {
          const int i = t + r * kRhThreads;
          if (i < S * A4) {
            const int s = i / A4, j = i - s * A4;
            reinterpret_cast<float4*>(L + (3 * s + d) * kRhArea)[j] = c[r];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();
      // inverse y: threads 0 .. 45 S - 1 take column q = 1 .. 15 of (slab s,
      // component d), the next 3 S threads the Hermitian columns 0 and 16 of
      // (s, d) packed into one complex FFT
      if (t < 45 * S) {
        const int sd = t / 15, q = 1 + t - 15 * sd;
        if (sd % 3 > 0 || need_x) {
          float2* col = L2 + sd * (kRhArea / 2) + q;
#pragma unroll
          for (int k = 0; k < G; ++k) v[k] = col[k * H];
          reg_fft<G, true>(v);
#pragma unroll
          for (int y = 0; y < G; ++y) col[y * H] = v[y];
        }
      } else if (t < 48 * S) {
        const int sd = t - 45 * S;
        if (sd % 3 > 0 || need_x) {
          float2* col = L2 + sd * (kRhArea / 2);
#pragma unroll
          for (int k = 0; k < G; ++k) {
            const float2 y0 = col[k * H], y16 = col[k * H + G / 2];
            v[k] = make_float2(y0.x - y16.y, y0.y + y16.x);
          }
          reg_fft<G, true>(v);
#pragma unroll
          for (int y = 0; y < G; ++y) {
            col[y * H] = make_float2(v[y].x, 0.f);
            col[y * H + G / 2] = make_float2(v[y].y, 0.f);
          }
        }
      }
      __syncthreads();
      const int s = t / 48, rem = t - 48 * s, d = rem >> 4, q = rem & 15;
      const bool act = s < S && (d > 0 || need_x);
      float2* const base = L2 + (3 * s + d) * (kRhArea / 2);
      if (act) {  // inverse x of rows q and q + 16
        float2* ra = base + q * H;
        float2* rb = base + (q + 16) * H;
#pragma unroll
        for (int k = 0; k <= G / 2; ++k) {
          const float2 A = ra[k], B = rb[k];
          v[k] = make_float2(A.x - B.y, A.y + B.x);
          if (k > 0 && k < G / 2) v[G - k] = make_float2(A.x + B.y, B.x - A.y);
        }
        reg_fft<G, true>(v);
        float* fa = reinterpret_cast<float*>(ra);
        float* fb = reinterpret_cast<float*>(rb);
#pragma unroll
        for (int x = 0; x < G; ++x) {
          fa[x] = v[x].x;
          fb[x] = v[x].y;
        }
      }
      __syncthreads();
      // pack: each thread takes pairs of neighbouring cells (x, x + 1), PB
      // pairs at a time (their LDS reads in flight together), and stores the
      // pair's two float4 as 32 contiguous bytes
      constexpr int NPK = (kRhRound * GG / 2 + kRhThreads - 1) / kRhThreads, PB = 2;
      const float r0x = Rr[0], r1x = Rr[1], r2x = Rr[2], r4x = Rr[4], r5x = Rr[5], r6x = Rr[6],
                  r8x = Rr[8], r9x = Rr[9], r10x = Rr[10];
      auto pack1 = [&](float fx, float fy, float fz) {
        const float mag = sqrtf(fmaf(fx, fx, fmaf(fy, fy, fz * fz)));
        if (world)
          return make_float4(fmaf(r0x, fx, fmaf(r1x, fy, r2x * fz)),
                             fmaf(r4x, fx, fmaf(r5x, fy, r6x * fz)),
                             fmaf(r8x, fx, fmaf(r9x, fy, r10x * fz)), mag);
        return make_float4(fx, fy, fz, mag);
      };
#pragma unroll
      for (int r0 = 0; r0 < NPK; r0 += PB) {
        float2 fx[PB], fy[PB], fz[PB];
#pragma unroll
        for (int r = 0; r < PB; ++r) {
          const int i = 2 * min(t + (r0 + r) * kRhThreads, S * GG / 2 - 1);
          const float* Fa = L + 3 * (i >> 10) * kRhArea + ((i >> 5) & (G - 1)) * kRhRow + (i & (G - 1));
          fy[r] = *reinterpret_cast<const float2*>(Fa + kRhArea);
          fz[r] = *reinterpret_cast<const float2*>(Fa + 2 * kRhArea);
          const float2 xv = *reinterpret_cast<const float2*>(Fa);  // (a value select, not a pointer one)
          fx[r] = need_x ? xv : fz[r];
        }
#pragma unroll
        for (int r = 0; r < PB; ++r) {
          const int i = 2 * (t + (r0 + r) * kRhThreads);
          if (r0 + r < NPK && i < S * GG) {
            float4* o = out + z0 * GG + i;
            o[0] = pack1(fx[r].x, fy[r].x, fz[r].x);
            o[1] = pack1(fx[r].y, fy[r].y, fz[r].y);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Workgroups of rh_grid_kernel for `ngrids` grids and `scratch_elems` float2
// of scratch: one per grid up to the CU count (one workgroup fits a CU) and
// the slots the scratch holds.
static long long rh_slots(long long ngrids, long long scratch_elems) {
  static int cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cus[dev] <= 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
      c = 256;
    cus[dev] = c;
  }
  return std::min(std::min(ngrids, (long long)cus[dev]), scratch_elems / kRhSlot);
}

long long rh_scratch_elems(long long ngrids) {
  int dev = 0, c = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
    c = 256;
  return std::min(ngrids, (long long)c) * kRhSlot;
}

template <typename R>
static hipError_t launch_rh_grids(const R& res, long long g0, long long ngrids, int world,
                                  float2* scratch, long long scratch_elems, hipStream_t s) {
  if (ngrids <= 0) return hipSuccess;
  const long long wgs = rh_slots(ngrids, scratch_elems);
  if (wgs <= 0) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void*)rh_grid_kernel<R>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRhLds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((rh_grid_kernel<R>), dim3((unsigned)wgs), dim3(kRhThreads), kRhLds, s, res,
                     g0, ngrids, world, scratch);
  return hipGetLastError();
}

// One rf_grid_kernel workgroup per grid, grids [g0, g0 + ngrids) of `res`.
template <int G, typename R>
static hipError_t launch_rf_grids(const R& res, long long g0, long long ngrids, int world,
                                  hipStream_t s) {
  if (ngrids <= 0) return hipSuccess;
  const size_t lds = (size_t)2 * G * G * (G + 1) * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)rf_grid_kernel<G, R>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((rf_grid_kernel<G, R>), dim3((unsigned)ngrids), dim3(512), lds, s, res, g0,
                     world);
  return hipGetLastError();
}

template <int G>
static hipError_t launch_rf_solve(const Params& P, const long long* grids, float2* S1,
                                  long long s1_elems, float2* S2, float4* force, const float* gtab,
                                  const float* kd, const float2* conv, int need_x, const float* rot,
                                  int world, hipStream_t s) {
  const RfSingle res = {grids, S1, S2, force, gtab, kd, conv, rot, need_x};
  if (G == 16) return launch_rf_grids<16>(res, 0, P.k, world, s);
  return launch_rh_grids(res, 0, P.k, world, S1, s1_elems, s);
}

// Two g = 16 solves of the same density grids in one launch of 2k
// workgroups: force_a from (gtab_a, conv_a), force_b from (gtab_b, conv_b).
// Other sizes return hipErrorNotSupported (the caller runs two solves).
hipError_t launch_solve_pair(const Params& P, const long long* grids, float4* force_a,
                             const float* gtab_a, const float2* conv_a, float4* force_b,
                             const float* gtab_b, const float2* conv_b, const float* kd,
                             int need_x, const float* rot, int world, hipStream_t s) {
  if (P.g != 16) return hipErrorNotSupported;
  const RfPair res = {{grids, nullptr, nullptr, force_a, gtab_a, kd, conv_a, rot, need_x},
                      {grids, nullptr, nullptr, force_b, gtab_b, kd, conv_b, rot, need_x},
                      (long long)P.k * 16};
  return launch_rf_grids<16>(res, 0, 2ll * P.k, world, s);
}

// ---------------------------------------------------------------------------
// Batched ensemble: the three passes, one launch each per grid-size class
// (g <= 16, g = 32), one workgroup per slab of any simulation of the class.
// A class lists its simulations (cls_sim) and their first slab in the
// class's flat slab range (cls_off, ascending); a slab runs the same body as
// the single-simulation launch, with the simulation's own g, tables and
// buffer offsets.  Splitting by class sizes LDS per class, so g = 16 slabs
// keep the occupancy of the single-simulation launch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int class_slot(const long long* __restrict__ cls_off, int ncls,
                                          long long t) {
  int lo = 0, hi = ncls - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cls_off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The spectral scratch of the batch covers one window of whole grids at a
// time: slab t of a window (slabs slab0 .. slab0 + nslabs of its class)
// belongs to class slot j, whose scratch starts at S + soff[j] (the window's
// offset of the simulation's grid 0, so that the slab bodies index the
// scratch as they index the simulation's own arrays).
__device__ RfSlab RfBatch::operator()(long long t) const {
  const int j = class_slot(cls_off, ncls, t);
  const BatchSim& b = sims[cls_sim[j]];
  const long long so = soff ? soff[j] : 0;  // (no scratch offsets: kernels without scratch)
  return {mass + b.cell_off,          S + so,
          S2 + so,                    force + b.cell_off,
          gtab + b.tab_off,           kd + b.line_off,
          b.conv_off >= 0 ? conv + b.conv_off : nullptr,
          rot + 12 * b.grid_off,      t - cls_off[j],
          b.need_x};
}

__global__ __launch_bounds__(256) void batch_fwd_xy_kernel(const BatchSim* __restrict__ sims,
                                                           const int* __restrict__ cls_sim,
                                                           const long long* __restrict__ cls_off,
                                                           int ncls, long long slab0,
                                                           const long long* __restrict__ soff,
                                                           const long long* __restrict__ mass,
                                                           float2* __restrict__ S,
                                                           const float2* __restrict__ tw) {
  const long long t = slab0 + blockIdx.x;
  const int j = class_slot(cls_off, ncls, t);
  const BatchSim& b = sims[cls_sim[j]];
  spec_fwd_xy_slab(mass + b.cell_off, S + soff[j], tw + b.line_off, b.P.g, b.P.lg,
                   t - cls_off[j]);
}

__global__ __launch_bounds__(256) void batch_z_green_kernel(const BatchSim* __restrict__ sims,
                                                            const int* __restrict__ cls_sim,
                                                            const long long* __restrict__ cls_off,
                                                            int ncls, long long slab0,
                                                            const long long* __restrict__ soff,
                                                            float2* __restrict__ S,
                                                            float2* __restrict__ S2,
                                                            const float2* __restrict__ tw,
                                                            const float* __restrict__ gtab,
                                                            const float* __restrict__ kd,
                                                            const float2* __restrict__ conv) {
  const long long t = slab0 + blockIdx.x;
  const int j = class_slot(cls_off, ncls, t);
  const BatchSim& b = sims[cls_sim[j]];
  spec_z_green_slab(S + soff[j], S2 + soff[j], tw + b.line_off, gtab + b.tab_off,
                    kd + b.line_off, b.conv_off >= 0 ? conv + b.conv_off : nullptr, b.P.g, b.P.lg,
                    b.need_x, t - cls_off[j]);
}

__global__ __launch_bounds__(256) void batch_inv_xy_kernel(const BatchSim* __restrict__ sims,
                                                           const int* __restrict__ cls_sim,
                                                           const long long* __restrict__ cls_off,
                                                           int ncls, long long slab0,
                                                           const long long* __restrict__ soff,
                                                           const float2* __restrict__ S,
                                                           const float2* __restrict__ S2,
                                                           float4* __restrict__ force,
                                                           const float2* __restrict__ tw,
                                                           const float* __restrict__ rot) {
  const long long t = slab0 + blockIdx.x;
  const int j = class_slot(cls_off, ncls, t);
  const BatchSim& b = sims[cls_sim[j]];
  spec_inv_xy_slab(S + soff[j], S2 + soff[j], force + b.cell_off, tw + b.line_off, b.P.g,
                   b.P.lg, b.need_x, rot + 12 * b.grid_off, 1, t - cls_off[j]);
}

// Dynamic LDS bytes of pass 0 (A), 1 (B) or 2 (C) for grids of size up to g
// (the passes take the FFT layout from kFftMinG on).
static size_t solve_lds(int g, int pass) {
  const size_t gg = (size_t)g * g;
  size_t n;
  if (g >= kFftMinG)
    n = g + (pass == 0 ? 2 * gg : 4 * gg);
  else
    n = g + (pass == 0 ? (gg + 1) / 2 + gg : pass == 1 ? 3 * gg : 4 * gg);
  return n * sizeof(float2);
}

hipError_t launch_batch_solve(const BatchSim* sims, const int* cls_sim, const long long* cls_off,
                              int ncls, long long slab0, long long slabs, const long long* soff,
                              int gmax, const long long* grids, float2* S1, long long s1_elems,
                              float2* S2, float4* force, const float2* tw, const float* gtab,
                              const float* kd, const float2* conv, const float* rot,
                              hipStream_t s) {
  if (slabs <= 0 || ncls <= 0) return hipSuccess;
  // a class of grid size 16 or 32 runs the single-simulation kernels (one
  // workgroup per grid, same arithmetic) over all of its grids at once
  if (gmax == 16 || gmax == 32) {
    const RfBatch res = {sims, cls_sim, cls_off, ncls, soff, grids, S1, S2, force, gtab, kd, conv, rot};
    // (slab0 and slabs are multiples of g: whole grids)
    return gmax == 16 ? launch_rf_grids<16>(res, slab0 / 16, slabs / 16, 1, s)
                      : launch_rh_grids(res, slab0 / 32, slabs / 32, 1, S1, s1_elems, s);
  }
  const size_t ldsA = solve_lds(gmax, 0), ldsB = solve_lds(gmax, 1), ldsC = solve_lds(gmax, 2);
  hipError_t e;
  if ((e = hipFuncSetAttribute((const void*)batch_fwd_xy_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)batch_z_green_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)batch_inv_xy_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  const dim3 grid((unsigned)slabs);
  hipLaunchKernelGGL(batch_fwd_xy_kernel, grid, dim3(256), ldsA, s, sims, cls_sim, cls_off, ncls,
                     slab0, soff, grids, S1, tw);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(batch_z_green_kernel, grid, dim3(256), ldsB, s, sims, cls_sim, cls_off, ncls,
                     slab0, soff, S1, S2, tw, gtab, kd, conv);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(batch_inv_xy_kernel, grid, dim3(256), ldsC, s, sims, cls_sim, cls_off, ncls,
                     slab0, soff, S1, S2, force, tw, rot);
  return hipGetLastError();
}

hipError_t launch_solve(const Params& P, const long long* grids, float2* S1, long long s1_elems,
                        float2* S2, float4* force, const float2* tw, const float* gtab,
                        const float* kd, const float2* conv, int need_x, const float* rot,
                        int world, hipStream_t s) {
  const int g = P.g, lg = P.lg;
  if (g == 16)
    return launch_rf_solve<16>(P, grids, S1, s1_elems, S2, force, gtab, kd, conv, need_x, rot,
                               world, s);
  if (g == 32)
    return launch_rf_solve<32>(P, grids, S1, s1_elems, S2, force, gtab, kd, conv, need_x, rot,
                               world, s);
  const unsigned slabs = (unsigned)(P.k * g);
  const size_t ldsA = solve_lds(g, 0), ldsB = solve_lds(g, 1), ldsC = solve_lds(g, 2);
  hipError_t e;
  if ((e = hipFuncSetAttribute((const void*)spec_fwd_xy_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)spec_z_green_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute((const void*)spec_inv_xy_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(spec_fwd_xy_kernel, dim3(slabs), dim3(256), ldsA, s, grids, S1, tw, g, lg);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(spec_z_green_kernel, dim3(slabs), dim3(256), ldsB, s, S1, S2, tw, gtab, kd,
                     conv, g, lg, need_x);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(spec_inv_xy_kernel, dim3(slabs), dim3(256), ldsC, s, S1, S2, force, tw, g,
                     lg, need_x, rot, world);
  return hipGetLastError();
}

}  // namespace mog
