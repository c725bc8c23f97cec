// Checks the host-side near-test thresholds of the runtime against their
// definitions, by brute force around each threshold:
//   sqrtf_le_threshold(r): T = max float f >= 0 with (double) sqrtf(f) <= r
//   sqrtf_lt_threshold(r): U = min float f >= 0 with (double) sqrtf(f) >= r
//   double_le_threshold(t): Td with (float) s <= t  <=>  s <= Td for all doubles s
// sqrtf is glibc's (correctly rounded, IEEE 754).  Prints "ok <cases>" or the
// first failure.
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

namespace abl {
float sqrtf_le_threshold(double r);
float sqrtf_lt_threshold(double r);
double double_le_threshold(float t);
}

static int fails = 0;
static void fail(const char *what, double a, double b) {
  if (fails++ < 5) std::printf("FAIL %s %.17g %.17g\n", what, a, b);
}

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> ur(0.0, 50.0);
  long cases = 0;
  double radii[] = {0.0, 1.0, 1.5, 2.0, 2.5, 5.0, 8.0, 10.0, 0.1, 0.035, 0.02, 0.12};
  for (int k = 0; k < 2000; k++) {
    const double r = k < 12 ? radii[k] : ur(rng);
    const float T = abl::sqrtf_le_threshold(r);
    const float U = abl::sqrtf_lt_threshold(r);
    // T: sqrtf(T) <= r < sqrtf(next(T)); U: sqrtf(prev(U)) < r <= sqrtf(U)
    if (!((double) sqrtf(T) <= r)) fail("T in", r, T);
    if (!((double) sqrtf(std::nextafter(T, HUGE_VALF)) > r)) fail("T max", r, T);
    if (U > 0.f && !((double) sqrtf(std::nextafter(U, -HUGE_VALF)) < r)) fail("U min", r, U);
    if (!((double) sqrtf(U) >= r)) fail("U in", r, U);
    cases += 4;
  }
  // double_le_threshold: every float near typical thresholds, doubles on both
  // sides of the rounding midpoints, plus the extremes
  std::uniform_real_distribution<float> uf(0.f, 2000.f);
  for (int k = 0; k < 20000; k++) {
    float t = k == 0 ? 0.f : k == 1 ? FLT_MAX : k == 2 ? 1.f : k == 3 ? -1.f : uf(rng);
    const double td = abl::double_le_threshold(t);
    if (!((float) td <= t)) fail("Td in", t, td);
    const double above = std::nextafter(td, HUGE_VAL);
    if ((float) above <= t) fail("Td max", t, td);
    // random doubles around the midpoints on either side of t
    const float up = std::nextafter(t, HUGE_VALF), dn = std::nextafter(t, -HUGE_VALF);
    const double mids[2] = {((double) t + (double) up) / 2, ((double) t + (double) dn) / 2};
    for (double m : mids) {
      if (!std::isfinite(m)) continue;
      double s = m;
      for (int j = 0; j < 4; j++) s = std::nextafter(s, -HUGE_VAL);
      for (int j = 0; j < 9; j++, s = std::nextafter(s, HUGE_VAL)) {
        if (((float) s <= t) != (s <= td)) fail("Td eq", t, s);
        cases++;
      }
    }
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("ok %ld\n", cases);
  return 0;
}
