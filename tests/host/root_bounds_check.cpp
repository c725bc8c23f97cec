// Host check of abl_root_bounds.hpp: for constants c, sqrtf(s) OP c against
// the bound comparisons, over every non-negative float s (exhaustive; argv[1]
// "full") or a sample around each bound plus random floats, and the special
// values (0, denormals, FLT_MAX, inf, NaN).  The host's sqrtf is the IEEE
// correctly rounded root.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "abl_root_bounds.hpp"

using namespace abl::rootb;

static unsigned long long bad = 0, checked = 0;

static void check(float s, double c) {
  const double r = (double) std::sqrt(s);  // float sqrtf, then exact widening
  const bool lt = r < c, le = r <= c, gt = r > c, ge = r >= c;
  checked++;
  if (lt != root_lt(s, c) || le != root_le(s, c) || gt != root_gt(s, c) || ge != root_ge(s, c)) {
    if (bad < 10)
      printf("mismatch s=%a c=%a: %d%d%d%d vs %d%d%d%d\n", (double) s, c, lt, le, gt, ge,
             root_lt(s, c), root_le(s, c), root_gt(s, c), root_ge(s, c));
    bad++;
  }
}

int main(int argc, char **argv) {
  const bool full = argc > 1 && !strcmp(argv[1], "full");
  // constants as models use them: radii, thresholds, double expressions,
  // exact floats, values between floats, zero, negatives, tiny, huge
  const double cs[] = {2.0, 10.0, 1.1 * 3.0, 0.5, 1.0, 0.1, 4.0 / 3.0, 1e-30, 1e-45,
                       3e38, 1e30, 0.0, -1.0, 7.0000001, (double) 2.0000002f,
                       (double) std::nextafter(2.0f, 3.0f), 65504.0, 16777217.0};
  std::mt19937_64 rng(7);
  for (double c : cs) {
    if (full && (c == 2.0 || c == 10.0 || c == 1.1 * 3.0)) {
      for (uint64_t u = 0; u <= 0x7F800000u; u++) {
        float s;
        uint32_t b = (uint32_t) u;
        memcpy(&s, &b, 4);
        check(s, c);
      }
    }
    // around the bounds
    const float bs[] = {root_bound(below(c)), root_bound(at_most(c))};
    for (float b0 : bs) {
      if (b0 < 0) continue;
      uint32_t b;
      memcpy(&b, &b0, 4);
      for (int d = -2000; d <= 2000; d++) {
        const int64_t v = (int64_t) b + d;
        if (v < 0 || v > 0x7F800000) continue;
        uint32_t u = (uint32_t) v;
        float s;
        memcpy(&s, &u, 4);
        check(s, c);
      }
    }
    for (int k = 0; k < 200000; k++) {
      uint32_t u = (uint32_t) (rng() % 0x7F800001u);
      float s;
      memcpy(&s, &u, 4);
      check(s, c);
    }
    const float sp[] = {0.f, -0.f, 1e-45f, 1e-38f, kFltMax, INFINITY, NAN};
    for (float s : sp) check(s, c);
  }
  printf("checked %llu bad %llu\n", checked, bad);
  return bad ? 1 : 0;
}
