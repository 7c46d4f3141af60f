// Host check of the kernels' workgroup-to-tile map (csrc/tiling.h).
//
// For every tile count T in [1, Tmax] (and a few large ones): the map is a
// bijection on [0, T); the tiles of each XCD (workgroups L with L % 8 = x)
// form one contiguous range in order of L; the ranges of XCDs 0..7 follow
// each other from 0 and differ in length by at most one.  Prints the number
// of failing T and exits non-zero on any failure.  Usage: tiling_check [Tmax]
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "tiling.h"

static bool check(unsigned int T) {
  std::vector<unsigned char> seen(T, 0);
  unsigned int next_start = 0, lo = T, hi = 0;
  for (unsigned int x = 0; x < 8; ++x) {
    unsigned int count = 0, prev = 0;
    for (unsigned int L = x; L < T; L += 8, ++count) {
      const unsigned int t = mog::xcd_tile(L, T);
      if (t >= T || seen[t]) return false;
      seen[t] = 1;
      if (count == 0 ? t != next_start : t != prev + 1) return false;
      prev = t;
    }
    next_start += count;
    if (count < lo) lo = count;
    if (count > hi) hi = count;
  }
  return next_start == T && hi - lo <= 1;
}

int main(int argc, char** argv) {
  const unsigned int tmax = argc > 1 ? (unsigned int)atoi(argv[1]) : 20000u;
  unsigned int bad = 0;
  for (unsigned int T = 1; T <= tmax; ++T) bad += !check(T);
  const unsigned int large[] = {4320u, 65536u + 3u, 1u << 20, (1u << 20) + 7u, 4000037u};
  for (unsigned int T : large) bad += !check(T);
  printf("%u\n", bad);
  return bad ? 1 : 0;
}
