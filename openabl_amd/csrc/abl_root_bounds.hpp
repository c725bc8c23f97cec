// abl_root_bounds.hpp -- comparisons of a correctly rounded single-precision
// square root with a constant, decided on the argument of the root.
//
// The C runtime's dist() and length() are (abl_float) sqrtf((float) s) for
// the float square sum s (asset/c/libabl.h:156-168).  A model that only
// compares such a distance with a constant, as boids' `separation <
// separation_radius` does, needs no root: sqrtf is monotone and correctly
// rounded, so for every float s and every constant c
//   sqrtf(s) <  c  <=>  s <= root_bound(below(c))
//   sqrtf(s) <= c  <=>  s <= root_bound(at_most(c))
//   sqrtf(s) >  c  <=>  s >  root_bound(at_most(c))
//   sqrtf(s) >= c  <=>  s >  root_bound(below(c))
// where below(c) / at_most(c) are the largest floats < c / <= c (the root is
// a float, and the comparison in double is exact), and root_bound(m) is the
// largest float whose rounded root is at most m: sqrtf(s) <= m exactly when
// sqrt(s) < m + h, h half the gap from m to the next float up (a root never
// falls on that midpoint: its square has 50 significant bits), i.e. when s
// is below (m + h)^2, which double holds exactly.  A NaN s fails every
// comparison on both sides.  Everything here is constexpr, so with a
// constant c the bound folds at compile time; no header of the GPU toolchain
// is needed, so the host test compiles it with g++.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define ABL_RB_FN __host__ __device__ constexpr inline
#else
#define ABL_RB_FN constexpr inline
#endif

namespace abl {
namespace rootb {

constexpr float kFltMax = 3.40282346638528859811704183484516925e+38f;

ABL_RB_FN uint32_t bits(float f) { return __builtin_bit_cast(uint32_t, f); }
ABL_RB_FN float from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

// the next float below a positive float (the largest negative float below 0
// is never needed: callers pass f > 0)
ABL_RB_FN float prev_pos(float f) { return from_bits(bits(f) - 1u); }

// largest float strictly below c (c finite; c > FLT_MAX gives FLT_MAX)
ABL_RB_FN float below(double c) {
  if (c > (double) kFltMax) return kFltMax;
  if (c < -(double) kFltMax) return -kFltMax;
  float f = (float) c;
  if ((double) f >= c) f = f > 0.f ? prev_pos(f) : (f == 0.f ? -1.f : from_bits(bits(f) + 1u));
  return f;
}

// largest float at most c
ABL_RB_FN float at_most(double c) {
  if (c >= (double) kFltMax) return kFltMax;
  if (c < -(double) kFltMax) return -kFltMax;
  float f = (float) c;
  if ((double) f > c) f = f > 0.f ? prev_pos(f) : (f == 0.f ? -1.f : from_bits(bits(f) + 1u));
  return f;
}

// largest float s with sqrtf(s) <= m; -1 when no s >= 0 qualifies (m < 0)
ABL_RB_FN float root_bound(float m) {
  if (m < 0.f) return -1.f;
  if (m >= kFltMax) return kFltMax;
  const float up = from_bits(bits(m) + 1u);  // next float above m (m >= +0)
  const double mid = (double) m + ((double) up - (double) m) * 0.5;
  const double b = mid * mid;  // exact: 25 bits squared
  if (b > (double) kFltMax) return kFltMax;
  float f = (float) b;
  if ((double) f >= b) f = prev_pos(f);
  return f;
}

// (s: the float whose correctly rounded root is the distance)
ABL_RB_FN bool root_lt(float s, double c) { return s <= root_bound(below(c)); }
ABL_RB_FN bool root_le(float s, double c) { return s <= root_bound(at_most(c)); }
ABL_RB_FN bool root_gt(float s, double c) { return s > root_bound(at_most(c)); }
ABL_RB_FN bool root_ge(float s, double c) { return s > root_bound(below(c)); }

}  // namespace rootb
}  // namespace abl
