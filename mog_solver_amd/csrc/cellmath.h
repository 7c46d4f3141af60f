// NGP cell arithmetic of the MOG step, shared by the device kernels and a
// host-side exhaustiveness check (plain C++, no HIP types).
//
// Exact path: the float32 sequence of mass.glsl:42-50 / update.glsl:71-76
//   c = p - W/2;  r_m = (R[m]*c0 + R[4+m]*c1) + R[8+m]*c2;  l = (r + W/2) - off
//   l = mod(mod(l, W) + W, W);  cell = int(l / h) % g
// with every operation rounded separately (the library is built with
// -ffp-contract=off) and IEEE division.
//
// Fast path: w_m = l_m / h - 1/2 evaluated as three FMAs with per-grid
// coefficients F[4m+j] = RN(R[4j+m] / h), F[4m+3] = RN((W/2 - off_m)/h - 1/2);
// the cell is floor(v) mod g = rint(w) mod g.  rint is taken by adding and
// subtracting 1.5 * 2^23: for |w| < 2^22 the sum t = w + 1.5 * 2^23 lies in
// [2^23, 2^24), where float32 spacing is 1, so t is w rounded to the
// nearest integer (ties to even, like rint), and the low bits of t's
// encoding are that integer modulo 2^23 (two's complement), from which the
// cell's coordinate is masked without a float-to-integer conversion; the
// face distance uses rint(w) itself.  Admitted particles give |w| < 8 g.  Both paths differ from the real-number value v* by at
// most a bound E (derived in fast_margin below); when v lies farther than
// M >= E_exact + E_fast from every integer, floor(v) mod g equals the exact
// path's cell, so the fast result is used; otherwise the exact sequence runs.
// The bound assumes |c_j| <= CB for the particle, which is checked per
// particle (non-finite coordinates fail the check and take the exact path).
#pragma once

#include <math.h>
#include <string.h>

#if defined(__HIPCC__)
#define MOG_HD __host__ __device__ __forceinline__
#else
#define MOG_HD inline
#endif

namespace mog {

// GLSL mod(x, y) = x - y * floor(x / y).
MOG_HD float glsl_mod_f(float x, float y) {
  float t = floorf(x / y);
  float p = y * t;
  return x - p;
}

MOG_HD int axis_cell_exact(float l, float W, float h, int gmask) {
  l = glsl_mod_f(l, W);
  l = l + W;
  l = glsl_mod_f(l, W);
  float q = l / h;
  int i = (int)q;
  return i & gmask;
}

// Exact cell from the centred position and the packed rotation row R
// (R[4j+m] = M[j][m]) and offset O of one grid.
MOG_HD int cell_exact(float c0, float c1, float c2, const float* R, const float* O, float W,
                      float half, float h, int lg) {
  float r0 = R[0] * c0;
  r0 = r0 + R[4] * c1;
  r0 = r0 + R[8] * c2;
  float r1 = R[1] * c0;
  r1 = r1 + R[5] * c1;
  r1 = r1 + R[9] * c2;
  float r2 = R[2] * c0;
  r2 = r2 + R[6] * c1;
  r2 = r2 + R[10] * c2;
  float l0 = (r0 + half) - O[0];
  float l1 = (r1 + half) - O[1];
  float l2 = (r2 + half) - O[2];
  const int gm = (1 << lg) - 1;
  int x = axis_cell_exact(l0, W, h, gm);
  int y = axis_cell_exact(l1, W, h, gm);
  int z = axis_cell_exact(l2, W, h, gm);
  return (((z << lg) | y) << lg) | x;
}

// Fast cell and its face distance: returns d = max_m |w_m - rint(w_m)|
// (the distance of v to the nearest integer is 0.5 - d on the worst axis)
// and writes the cell floor(v) mod g.  F is the 12-float fast row of the
// grid (make_fast_row) with its three addends F[3], F[7], F[11] passed as
// values (a kernel can then keep them in vector registers).  The cell is in
// range for every input; d is meaningful for admitted particles only.
MOG_HD float cell_fast_d(float c0, float c1, float c2, const float* F, float a0, float a1, float a2,
                         int lg, int* cell) {
  const float w0 = fmaf(F[2], c2, fmaf(F[1], c1, fmaf(F[0], c0, a0)));
  const float w1 = fmaf(F[6], c2, fmaf(F[5], c1, fmaf(F[4], c0, a1)));
  const float w2 = fmaf(F[10], c2, fmaf(F[9], c1, fmaf(F[8], c0, a2)));
  const float M = 12582912.0f;  // 1.5 * 2^23
  const float t0 = w0 + M, t1 = w1 + M, t2 = w2 + M;
  // rint(w) == t - M for |w| < 2^22; rintf keeps the two results independent
  const float r0 = rintf(w0), r1 = rintf(w1), r2 = rintf(w2);
  unsigned b0, b1, b2;
  memcpy(&b0, &t0, 4);
  memcpy(&b1, &t1, 4);
  memcpy(&b2, &t2, 4);
  // bit-field inserts: x from b0, y from b1 << lg, z from b2 << 2 lg
  const unsigned gm = (1u << lg) - 1u, gm2 = (1u << (2 * lg)) - 1u;
  unsigned c = (b0 & gm) | ((b1 << lg) & ~gm);
  c = (c & gm2) | ((b2 << (2 * lg)) & ~gm2);
  *cell = (int)(c & ((1u << (3 * lg)) - 1u));
  return fmaxf(fmaxf(fabsf(w0 - r0), fabsf(w1 - r1)), fabsf(w2 - r2));
}

// Shifted cell fields.  For an integer r with |r| < 2^(22 - s), the float32
// sum u = r + 1.5 * 2^(23 - s) is exact (u lies in [2^(23-s), 2^(24-s)),
// where the spacing is 2^-s), and the low 23 bits of its encoding are
// r * 2^s + 2^22: the integer sits at bit s (two's complement, zeros
// below).  With s_m = s0 + m lg for the three axes the coordinates of the
// cell land in disjoint fields, and two bit selects and a mask give
// cell << s0: the byte offset of the cell in an array of 2^s0-byte
// elements.  The adds and bit selects issue at the full vector rate on
// gfx950, where shifts and shift-adds take twice as long.
struct FieldShift {
  float k0, k1, k2;         // 1.5 * 2^(23 - s_m)
  unsigned m0, m01, m012;   // bits below s0 + lg, s0 + 2 lg; bits [s0, s0 + 3 lg)
};

MOG_HD float pow2_1p5(int e) {  // 1.5 * 2^e, |e| < 127
  const unsigned b = ((unsigned)(127 + e) << 23) | 0x400000u;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

MOG_HD FieldShift field_shift(int lg, int s0) {
  FieldShift S;
  S.k0 = pow2_1p5(23 - s0);
  S.k1 = pow2_1p5(23 - s0 - lg);
  S.k2 = pow2_1p5(23 - s0 - 2 * lg);
  S.m0 = (1u << (s0 + lg)) - 1u;
  S.m01 = (1u << (s0 + 2 * lg)) - 1u;
  // the final mask also clears the bits below s0: zero for admitted
  // particles, they hold rounding residue where |w| exceeds the fields'
  // range (a particle outside the admitted bound), and the offset must stay
  // a multiple of the element size for every input
  S.m012 = ((1u << (s0 + 3 * lg)) - 1u) & ~((1u << s0) - 1u);
  return S;
}

// Bound on |w_m| of the fast path for admitted particles (|c_j| <= CB) of
// transforms with |R| <= rmax, |off| <= omax: |F| <= rmax / h and
// |F[4m+3]| <= (W/2 + omax) / h + 1/2, each up to one rounding, and three
// FMAs of relative error 2^-24 each; 1.0 covers the 1/2 and the slack.
inline double fast_w_bound(double W, int g, double rmax, double omax, double CB) {
  const double h = W / g;
  return (3.0 * rmax * CB + 0.5 * W + omax) / h * (1.0 + 1e-6) + 1.0;
}

// Whether the shifted fields hold for every admitted particle: |rint(w)| <=
// wbound + 1/2 must stay below 2^(22 - s0 - 2 lg), and the three fields
// below bit 22.
inline bool field_shift_ok(int lg, int s0, double wbound) {
  return s0 + 3 * lg <= 22 && wbound + 0.5 < ldexp(1.0, 22 - s0 - 2 * lg);
}

// Bit operations with three operands, one v_bitop3_b32 each on the device
// (truth-table immediates: operands a, b, c select table bits 0xf0, 0xcc,
// 0xaa).  bit_select(a, b, m) = (a & m) | (b & ~m); and_or(a, m, c) =
// (a & m) | c.
MOG_HD unsigned bit_select(unsigned a, unsigned b, unsigned m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, m, 0xE4);
#else
  return (a & m) | (b & ~m);
#endif
}

MOG_HD unsigned and_or(unsigned a, unsigned m, unsigned c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, m, c, 0xEA);
#else
  return (a & m) | c;
#endif
}

// cell_fast_d with the cell returned as the byte offset (cell << s0) | base
// (S from field_shift(lg, s0), valid under field_shift_ok; base has no bits
// below s0 + 3 lg).  Same cell, same d, for admitted particles; for every
// input the offset is a multiple of 2^s0 below g^3 << s0 (an aligned element
// inside the grid), whatever the cell.
MOG_HD float cell_fast_o(float c0, float c1, float c2, const float* F, float a0, float a1, float a2,
                         const FieldShift& S, unsigned base, unsigned* off) {
  const float w0 = fmaf(F[2], c2, fmaf(F[1], c1, fmaf(F[0], c0, a0)));
  const float w1 = fmaf(F[6], c2, fmaf(F[5], c1, fmaf(F[4], c0, a1)));
  const float w2 = fmaf(F[10], c2, fmaf(F[9], c1, fmaf(F[8], c0, a2)));
  const float r0 = rintf(w0), r1 = rintf(w1), r2 = rintf(w2);
  const float u0 = r0 + S.k0, u1 = r1 + S.k1, u2 = r2 + S.k2;
  unsigned b0, b1, b2;
  memcpy(&b0, &u0, 4);
  memcpy(&b1, &u1, 4);
  memcpy(&b2, &u2, 4);
  const unsigned x = bit_select(bit_select(b0, b1, S.m0), b2, S.m01);
  *off = and_or(x, S.m012, base);
  return fmaxf(fmaxf(fabsf(w0 - r0), fabsf(w1 - r1)), fabsf(w2 - r2));
}

// Fast cell; returns 0 when the exact path must decide (lim = 0.5 - M).
// With w = v - 0.5 (the -0.5 is folded into F[4m+3]) and r = rint(w):
// floor(v) = r.  Callers admit only finite coordinates, so no NaN reaches
// the comparison.
MOG_HD int cell_fast_a(float c0, float c1, float c2, const float* F, float a0, float a1, float a2,
                       float lim, int lg, int* cell) {
  return (int)(cell_fast_d(c0, c1, c2, F, a0, a1, a2, lg, cell) < lim);
}

MOG_HD int cell_fast(float c0, float c1, float c2, const float* F, float lim, int lg, int* cell) {
  return cell_fast_a(c0, c1, c2, F, F[3], F[7], F[11], lim, lg, cell);
}

// The 12-float fast row of one grid: F[4m+j] = RN(R[4j+m] / h),
// F[4m+3] = RN((W/2 - off_m) / h - 1/2), from the float32 values the
// kernels use (half = float(W) * 0.5f, h = float(W / g)).
inline void make_fast_row(const float* R, const float* O, float half, float h, float* F) {
  for (int m = 0; m < 3; ++m) {
    for (int j = 0; j < 3; ++j) F[4 * m + j] = (float)((double)R[4 * j + m] / (double)h);
    F[4 * m + 3] = (float)(((double)half - (double)O[m]) / (double)h - 0.5);
  }
}

// Per-particle admission to the fast path: every |c_j| <= CB (NaN fails).
MOG_HD int particle_fast(float c0, float c1, float c2, float CB) {
  return (int)(fabsf(c0) <= CB) & (int)(fabsf(c1) <= CB) & (int)(fabsf(c2) <= CB);
}

// Error margin M of the fast path (in cell units) for transforms whose
// entries satisfy |R| <= rmax and |off| <= omax, particles admitted with
// |c_j| <= CB, cell size h = W / g.  With u = 2^-24 and T = rmax*CB:
//   exact path, l: products, partial sums, +W/2, -off     u (14 T + W + omax)
//   exact path, mod chain (|floor| <= L/W, L = 3T + W/2 + omax)  u (L + 4 W)
//   exact path, l / h                                   u g      (cell units)
//   fast path, coefficient rounding                     u (3 T + W/2 + omax + h/2)
//   fast path, three FMAs                               u (6 T + 3 (W/2 + omax + h/2))
// M = 3 x (sum / h + u g); returns a value >= 0.05 when the fast path should
// not be used.
inline double fast_margin(double W, int g, double rmax, double omax, double CB) {
  const double u = 1.0 / 16777216.0;
  const double h = W / g;
  const double T = rmax * CB;
  const double L = 3.0 * T + 0.5 * W + omax;
  const double exact_l = 14.0 * T + W + omax;
  const double exact_mod = L + 4.0 * W;
  const double fast = 3.0 * T + 0.5 * W + omax + 0.5 * h + 6.0 * T + 3.0 * (0.5 * W + omax + 0.5 * h);
  return 3.0 * (u * (exact_l + exact_mod + fast) / h + u * g);
}

}  // namespace mog
