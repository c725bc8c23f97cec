/* div_recip_check.c -- the shared-reciprocal quotient of the step programs
 * (div_by_recip with the div_ok range, openabl_amd/csrc/abl_device.hpp),
 * restated in C, against IEEE division: q = a*y, r = fma(-b, q, a),
 * q' = fma(r, y, q) with y = RN(1/b) must equal a/b bit for bit whenever
 * |a| and |b| lie in [2^-400, 2^400] (double) or [2^-50, 2^50] (float).
 * Cases: counts 1..255 as divisors with numerators of every sign and a
 * spread of exponents; random divisors over the whole range; numerators a
 * few ulps from b times a quotient (quotients next to rounding ties); the
 * force quotients of the circle model ((q - x) * s / |q - x| with a float
 * distance); the range ends.  Prints the number of mismatches (0 expected)
 * per kind and exits non-zero on any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) { /* splitmix64 */
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double u01(void) { return (double) (next() >> 11) * 0x1p-53; }

static double rand_d(int emin, int emax) {
  double m = 1.0 + u01();
  int e = emin + (int) (next() % (uint64_t) (emax - emin + 1));
  double v = ldexp(m, e);
  return (next() & 1) ? -v : v;
}
static float rand_f(int emin, int emax) {
  float m = 1.0f + (float) (next() >> 40) * 0x1p-24f;
  int e = emin + (int) (next() % (uint64_t) (emax - emin + 1));
  float v = ldexpf(m, e);
  return (next() & 1) ? -v : v;
}

static double q_d(double a, double b, double y) {
  double q = a * y;
  double r = fma(-b, q, a);
  return fma(r, y, q);
}
static float q_f(float a, float b, float y) {
  float q = a * y;
  float r = fmaf(-b, q, a);
  return fmaf(r, y, q);
}
static int ok_d(double x) { return fabs(x) >= 0x1p-400 && fabs(x) <= 0x1p400; }
static int ok_f(float x) { return fabsf(x) >= 0x1p-50f && fabsf(x) <= 0x1p50f; }
static int same_d(double x, double y) { return memcmp(&x, &y, 8) == 0; }
static int same_f(float x, float y) { return memcmp(&x, &y, 4) == 0; }

int main(int argc, char **argv) {
  long n = argc > 1 ? atol(argv[1]) : 2000000;
  long bad[8] = {0};
  for (long i = 0; i < n; i++) {
    int c = 1 + (int) (next() % 255);
    double a = rand_d(-60, 60), b = (double) c;
    if (!same_d(q_d(a, b, 1.0 / b), a / b)) bad[0]++;
    float af = rand_f(-40, 40), bf = (float) c;
    if (!same_f(q_f(af, bf, 1.0f / bf), af / bf)) bad[1]++;
    double b2 = rand_d(-400, 399), a2 = rand_d(-400, 399);
    if (!same_d(q_d(a2, b2, 1.0 / b2), a2 / b2)) bad[2]++;
    float b2f = rand_f(-50, 49), a2f = rand_f(-50, 49);
    if (!same_f(q_f(a2f, b2f, 1.0f / b2f), a2f / b2f)) bad[3]++;
    double qv = rand_d(-30, 30), b3 = rand_d(-20, 20), a3 = b3 * qv;
    for (int k = 0; k < 4; k++) a3 = nextafter(a3, (next() & 1) ? INFINITY : -INFINITY);
    if (ok_d(a3) && !same_d(q_d(a3, b3, 1.0 / b3), a3 / b3)) bad[4]++;
    float qf = rand_f(-20, 20), b3f = rand_f(-15, 15), a3f = b3f * qf;
    for (int k = 0; k < 4; k++) a3f = nextafterf(a3f, (next() & 1) ? INFINITY : -INFINITY);
    if (ok_f(a3f) && !same_f(q_f(a3f, b3f, 1.0f / b3f), a3f / b3f)) bad[5]++;
    /* circle force: pos_dist = (double) sqrtf((float) s), s the square sum */
    double dx = rand_d(-6, 2), dy = rand_d(-6, 2), dz = rand_d(-6, 2);
    double pd = (double) sqrtf((float) (dx * dx + dy * dy + dz * dz));
    double sc = 0.01 * (8.0 - pd);
    double num = dx * sc;
    if (ok_d(num) && ok_d(pd) && !same_d(q_d(num, pd, 1.0 / pd), num / pd)) bad[6]++;
  }
  const double ends[] = {0x1p-400, 0x1.fffffffffffffp-400, 0x1p400, 0x1.fffffffffffffp399};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double a = ends[i], b = ends[j];
      if (!same_d(q_d(a, b, 1.0 / b), a / b)) bad[7]++;
      if (!same_d(q_d(-a, b, 1.0 / b), -a / b)) bad[7]++;
    }
  printf("counts_f64=%ld counts_f32=%ld random_f64=%ld random_f32=%ld near_tie_f64=%ld "
         "near_tie_f32=%ld circle_force=%ld range_ends=%ld\n",
         bad[0], bad[1], bad[2], bad[3], bad[4], bad[5], bad[6], bad[7]);
  long total = 0;
  for (int k = 0; k < 8; k++) total += bad[k];
  return total != 0;
}
