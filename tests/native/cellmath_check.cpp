// Host check of the fast NGP cell path against the exact float32 sequence.
//
// For random rotations/offsets, both admission bounds the library uses
// (|c_j| <= W/2 for the 3D step, 1.5 W for the planar path), and positions
// drawn both uniformly and adversarially on cell faces (a local coordinate
// j*h + tiny, mapped back to world space and rounded to float32), every
// position the fast path accepts must give the exact path's cell.  Every admitted position must also give
// the same cell through the shifted fields (cell_fast_o, byte offsets for
// 4- and 16-byte elements) wherever field_shift_ok admits them.  Prints
// "mismatches fastrate total shifted_checked" and exits non-zero on any
// mismatch.  Usage: cellmath_check [samples] [margin_scale]
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "cellmath.h"

static uint64_t s_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uni() { return (next_u64() >> 11) * (1.0 / 9007199254740992.0); }

struct Grid {
  float R[12], O[4], F[12];
  double M[3][3];
};

static void make_grid(Grid& G, double W, int g) {
  double q[4], n = 0;
  for (int i = 0; i < 4; ++i) {
    q[i] = uni() * 2 - 1;
    n += q[i] * q[i];
  }
  n = sqrt(n);
  for (int i = 0; i < 4; ++i) q[i] /= n;
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double M[3][3] = {{1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w},
                    {2 * x * y + 2 * z * w, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w},
                    {2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x * x - 2 * y * y}};
  const float h = (float)(W / g);
  const float half = (float)W * 0.5f;
  for (int j = 0; j < 3; ++j) {
    for (int m = 0; m < 3; ++m) {
      G.R[4 * j + m] = (float)M[j][m];
      G.M[j][m] = G.R[4 * j + m];
    }
    G.R[4 * j + 3] = 0;
    G.O[j] = (float)(uni() * h);
  }
  G.O[3] = 0;
  mog::make_fast_row(G.R, G.O, half, h, G.F);
}

int main(int argc, char** argv) {
  long long samples = argc > 1 ? atoll(argv[1]) : 2000000;
  // argv[2] scales the margin; 0 removes it (the check must then fail).
  const double mscale = argc > 2 ? atof(argv[2]) : 1.0;
  const double Ws[] = {100.0, 2000.0, 10.0, 1.0, 37.3};
  const int gs[] = {2, 4, 8, 16, 32, 64};
  long long bad = 0, fast = 0, total = 0, shifted = 0;
  const double cbws[] = {0.5, 1.5};
  for (double cbw : cbws)
  for (double W : Ws)
    for (int g : gs) {
      int lg = 0;
      while ((1 << lg) < g) ++lg;
      const float Wf = (float)W, half = Wf * 0.5f, h = (float)(W / g);
      const float CB = (float)(cbw * W);
      for (int t = 0; t < 8; ++t) {
        Grid G;
        make_grid(G, W, g);
        double rmax = 0, omax = 0;
        for (int i = 0; i < 12; ++i) rmax = fmax(rmax, fabs((double)G.R[i]));
        for (int i = 0; i < 3; ++i) omax = fmax(omax, fabs((double)G.O[i]));
        const double M = mog::fast_margin(W, g, rmax, omax, (double)CB) * mscale;
        const float lim = (float)(0.5 - M);
        const double wb = mog::fast_w_bound(W, g, rmax, omax, (double)CB);
        const bool sh2 = mog::field_shift_ok(lg, 2, wb), sh4 = mog::field_shift_ok(lg, 4, wb);
        const mog::FieldShift S2 = mog::field_shift(lg, 2), S4 = mog::field_shift(lg, 4);
        for (long long s = 0; s < samples / 240; ++s) {
          float p[3];
          if (s % 16 == 3) {  // far outside the box, up to 1e10 W
            for (int d = 0; d < 3; ++d) p[d] = (float)((uni() * 2.0 - 1.0) * W * pow(10.0, uni() * 10.0));
          } else if (s & 1) {
            for (int d = 0; d < 3; ++d) p[d] = (float)((uni() * 2.0 * cbw + 0.5 - cbw) * W * 1.25);
          } else {
            // local coordinates on or next to cell faces, mapped to world
            double loc[3];
            for (int d = 0; d < 3; ++d) {
              const int j = (int)(uni() * (g + 2)) - 1;
              const double tiny = (uni() - 0.5) * 4e-5 * h * (1 + (next_u64() % 4) * 10);
              loc[d] = j * (double)h + tiny;
            }
            for (int j = 0; j < 3; ++j) {
              double cj = 0;
              for (int m = 0; m < 3; ++m) cj += G.M[j][m] * (loc[m] - 0.5 * W + G.O[m]);
              p[j] = (float)(cj + 0.5 * W);
            }
          }
          const float c0 = p[0] - half, c1 = p[1] - half, c2 = p[2] - half;
          const int ce = mog::cell_exact(c0, c1, c2, G.R, G.O, Wf, half, h, lg);
          int cf = -1;
          const int ok = mog::cell_fast(c0, c1, c2, G.F, lim, lg, &cf) & mog::particle_fast(c0, c1, c2, CB);
          ++total;
          // every position, admitted or not: aligned offsets inside the grid
          for (int v = 0; v < 2; ++v) {
            const int s0 = v ? 4 : 2;
            unsigned off = ~0u;
            (void)mog::cell_fast_o(c0, c1, c2, G.F, G.F[3], G.F[7], G.F[11], v ? S4 : S2, 0u, &off);
            if ((off & ((1u << s0) - 1u)) != 0u || off >= (1u << (s0 + 3 * lg))) {
              if (bad < 10)
                fprintf(stderr, "offset out of range W=%g g=%d s0=%d p=(%.9g,%.9g,%.9g) off=%u\n", W, g, s0, p[0], p[1],
                        p[2], off);
              ++bad;
            }
          }
          if (mog::particle_fast(c0, c1, c2, CB)) {
            const float dref = mog::cell_fast_d(c0, c1, c2, G.F, G.F[3], G.F[7], G.F[11], lg, &cf);
            for (int v = 0; v < 2; ++v) {
              if (!(v ? sh4 : sh2)) continue;
              const int s0 = v ? 4 : 2;
              unsigned off = ~0u;
              const unsigned base = (unsigned)(t + 1) << (s0 + 3 * lg);
              const float d = mog::cell_fast_o(c0, c1, c2, G.F, G.F[3], G.F[7], G.F[11], v ? S4 : S2, base, &off);
              ++shifted;
              if (off != (((unsigned)cf << s0) | base) || d != dref) {
                if (bad < 10)
                  fprintf(stderr, "shifted mismatch W=%g g=%d s0=%d p=(%.9g,%.9g,%.9g) cell=%d off=%u\n", W, g,
                          s0, p[0], p[1], p[2], cf, off);
                ++bad;
              }
            }
          }
          if (ok) {
            ++fast;
            if (cf != ce) {
              if (bad < 10)
                fprintf(stderr, "mismatch W=%g g=%d p=(%.9g,%.9g,%.9g) exact=%d fast=%d\n", W, g, p[0], p[1],
                        p[2], ce, cf);
              ++bad;
            }
          }
        }
      }
    }
  printf("%lld %.6f %lld %lld\n", bad, (double)fast / (double)total, total, shifted);
  return bad ? 1 : 0;
}
