// abl_host.cpp -- host half of the -b hip runtime: the generator, the JSON
// writer and the near-test thresholds.  Compiled with the host C++ compiler.
//
// The generator and writer reproduce asset/c/libabl.c so that host-side
// initialisation and save() of a -b hip program produce exactly the bytes the
// C backend produces:
//   * xorshift128+ from {0xdeadbeef, 0xbeefdead} (libabl.c:8-17);
//   * random_float(lo, hi) = lo + (F) x / (F) (2^64 / (hi - lo)) evaluated in
//     the build's float type F (libabl.c:19-24);
//   * random_int: for a power-of-two span returns x & (n-1) WITHOUT adding the
//     lower bound; otherwise rejection sampling on the low 32 bits
//     (libabl.c:26-39);
//   * save(): "%f" members, [x,y(,z)] vectors, true/false, ",\n" between
//     agents, types and members in declaration order (libabl.c:46-124).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <mutex>

#include "../../include/abl_hip.h"

namespace {

uint64_t g_rng[2] = {0xdeadbeefULL, 0xbeefdeadULL};
std::mutex g_rng_mu;

uint64_t next_bits() {
  uint64_t a = g_rng[0];
  const uint64_t b = g_rng[1];
  g_rng[0] = b;
  a ^= a << 23;
  g_rng[1] = a ^ b ^ (a >> 17) ^ (b >> 26);
  return g_rng[1] + b;
}

}  // namespace

extern "C" double abl_random_float(double lo, double hi) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  uint64_t x = next_bits();
  double span = (double) UINT64_MAX / (hi - lo);
  return lo + (double) x / span;
}

extern "C" float abl_random_floatf(float lo, float hi) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  uint64_t x = next_bits();
  float span = (float) UINT64_MAX / (hi - lo);
  return lo + (float) x / span;
}

extern "C" int abl_random_int(int lo, int hi) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  unsigned n = (unsigned) (hi - lo + 1);
  if ((n & (n - 1)) == 0) return (int) (next_bits() & (n - 1));
  const unsigned limit = UINT_MAX - UINT_MAX % n;
  unsigned x;
  do {
    x = (unsigned) next_bits();
  } while (x >= limit);
  return lo + (int) (x % n);
}

extern "C" void abl_rng_seed(uint64_t s0, uint64_t s1) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  g_rng[0] = s0;
  g_rng[1] = s1;
}

extern "C" void abl_rng_state(uint64_t out[2]) {
  std::lock_guard<std::mutex> g(g_rng_mu);
  out[0] = g_rng[0];
  out[1] = g_rng[1];
}

// ---- JSON ------------------------------------------------------------------

namespace {

void write_value(FILE *f, const char *p, abl_type_id t, int fb) {
  auto rd = [fb](const char *q, int k) -> double {
    return fb == 8 ? ((const double *) q)[k] : (double) ((const float *) q)[k];
  };
  switch (t) {
    case ABL_TYPE_BOOL: fputs(*(const bool *) p ? "true" : "false", f); break;
    case ABL_TYPE_INT: fprintf(f, "%d", *(const int *) p); break;
    case ABL_TYPE_FLOAT: fprintf(f, "%f", rd(p, 0)); break;
    case ABL_TYPE_FLOAT2: fprintf(f, "[%f,%f]", rd(p, 0), rd(p, 1)); break;
    case ABL_TYPE_FLOAT3: fprintf(f, "[%f,%f,%f]", rd(p, 0), rd(p, 1), rd(p, 2)); break;
    default: break;
  }
}

}  // namespace

extern "C" int abl_save_json(const void *agents, const abl_agent_info *info, const char *path,
                             int float_bytes) {
  if (!agents || !info || !path || (float_bytes != 4 && float_bytes != 8)) return ABL_ERR_ARG;
  FILE *f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "save(): Count not open \"%s\" for writing\n", path);
    return ABL_ERR_ARG;
  }
  fputc('{', f);
  for (const abl_agent_info *a = info; a->name; a++) {
    if (a != info) fputc(',', f);
    fprintf(f, "\"%s\":", a->name);
    const abl_dyn_array *arr = (const abl_dyn_array *) ((const char *) agents + a->offset);
    // the terminator's offset is sizeof(struct)
    const abl_type_info *end = a->info;
    while (end->type != ABL_TYPE_END) end++;
    const size_t rec = end->offset;
    fputc('[', f);
    for (size_t i = 0; i < arr->len; i++) {
      if (i) fputs(",\n", f);
      const char *r = (const char *) arr->values + i * rec;
      fputc('{', f);
      for (const abl_type_info *m = a->info; m->type != ABL_TYPE_END; m++) {
        if (m != a->info) fputc(',', f);
        fprintf(f, "\"%s\":", m->name);
        write_value(f, r + m->offset, m->type, float_bytes);
      }
      fputc('}', f);
    }
    fputc(']', f);
  }
  fputc('}', f);
  fclose(f);
  return ABL_OK;
}

// ---- FLAME / FlameGPU XML ----------------------------------------------------
// save(..., SAVE_FLAME_XML / SAVE_FLAMEGPU_XML) (asset/c/libabl.c:126-213):
// <states>, <itno>0</itno>, one <xagent> per agent with its type's <name>
// and one element per member (vectors as name_x, name_y[, name_z]); FlameGPU
// writes the position member as x, y, z (z = 0.0 in 2-D).  Bools print as
// 0/1 (the reference reads the bool's bytes as an int, which is 0/1 while
// the struct padding after it is zero, as in every generated struct).

namespace {

void write_xml_member(FILE *f, const char *p, const abl_type_info *m, int fb, bool for_gpu) {
  auto rd = [fb](const char *q, int k) -> double {
    return fb == 8 ? ((const double *) q)[k] : (double) ((const float *) q)[k];
  };
  const char *n = m->name;
  if (for_gpu && m->is_pos) {
    if (m->type == ABL_TYPE_FLOAT2) {
      fprintf(f, "<x>%f</x>\n<y>%f</y>\n<z>0.0</z>\n", rd(p, 0), rd(p, 1));
    } else {
      fprintf(f, "<x>%f</x>\n<y>%f</y>\n<z>%f</z>\n", rd(p, 0), rd(p, 1), rd(p, 2));
    }
    return;
  }
  switch (m->type) {
    case ABL_TYPE_BOOL: fprintf(f, "<%s>%d</%s>\n", n, *(const bool *) p ? 1 : 0, n); break;
    case ABL_TYPE_INT: fprintf(f, "<%s>%d</%s>\n", n, *(const int *) p, n); break;
    case ABL_TYPE_FLOAT: fprintf(f, "<%s>%f</%s>\n", n, rd(p, 0), n); break;
    case ABL_TYPE_FLOAT2:
      fprintf(f, "<%s_x>%f</%s_x>\n<%s_y>%f</%s_y>\n", n, rd(p, 0), n, n, rd(p, 1), n);
      break;
    case ABL_TYPE_FLOAT3:
      fprintf(f, "<%s_x>%f</%s_x>\n<%s_y>%f</%s_y>\n<%s_z>%f</%s_z>\n", n, rd(p, 0), n, n,
              rd(p, 1), n, n, rd(p, 2), n);
      break;
    default: break;
  }
}

}  // namespace

extern "C" int abl_save_flame_xml(const void *agents, const abl_agent_info *info, const char *path,
                                  int float_bytes, int for_gpu) {
  if (!agents || !info || !path || (float_bytes != 4 && float_bytes != 8)) return ABL_ERR_ARG;
  for (const abl_agent_info *a = info; a->name; a++) {
    for (const abl_type_info *m = a->info; m->type != ABL_TYPE_END; m++) {
      if (for_gpu && m->is_pos && m->type != ABL_TYPE_FLOAT2 && m->type != ABL_TYPE_FLOAT3)
        return ABL_ERR_ARG;  // FlameGPU positions are vectors
    }
  }
  FILE *f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "save(): Count not open \"%s\" for writing\n", path);
    return ABL_ERR_ARG;
  }
  fputs("<states>\n<itno>0</itno>\n", f);
  for (const abl_agent_info *a = info; a->name; a++) {
    const abl_dyn_array *arr = (const abl_dyn_array *) ((const char *) agents + a->offset);
    const abl_type_info *end = a->info;
    while (end->type != ABL_TYPE_END) end++;
    const size_t rec = end->offset;
    for (size_t i = 0; i < arr->len; i++) {
      const char *r = (const char *) arr->values + i * rec;
      fprintf(f, "<xagent>\n<name>%s</name>\n", a->name);
      for (const abl_type_info *m = a->info; m->type != ABL_TYPE_END; m++)
        write_xml_member(f, r + m->offset, m, float_bytes, for_gpu != 0);
      fputs("</xagent>\n", f);
    }
  }
  fputs("</states>\n", f);
  fclose(f);
  return ABL_OK;
}

// ---- near-test thresholds --------------------------------------------------

namespace abl {

static float bits_to_float(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}

// Largest float f >= 0 with (double) sqrtf(f) <= r, or -1 if none.
float sqrtf_le_threshold(double r) {
  if (!(r >= 0.0)) return -1.0f;
  uint32_t lo = 0, hi = 0x7f800000u;  // +0 .. +inf
  if ((double) sqrtf(bits_to_float(hi)) <= r) return bits_to_float(hi);
  // invariant: sqrtf(lo) <= r < sqrtf(hi)
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if ((double) sqrtf(bits_to_float(mid)) <= r) lo = mid; else hi = mid;
  }
  return bits_to_float(lo);
}

// Smallest float f >= 0 with (double) sqrtf(f) >= r (so sqrtf(f) < r <=> f < U).
float sqrtf_lt_threshold(double r) {
  if (!(r > 0.0)) return 0.0f;
  uint32_t lo = 0, hi = 0x7f800000u;
  if ((double) sqrtf(bits_to_float(hi)) < r) return bits_to_float(hi);
  // invariant: sqrtf(lo) < r <= sqrtf(hi)
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if ((double) sqrtf(bits_to_float(mid)) < r) lo = mid; else hi = mid;
  }
  return bits_to_float(hi);
}

// Largest double d with (float) d <= t under round-to-nearest-even, so that
// for every double s:  (float) s <= t  <=>  s <= double_le_threshold(t).
// Between t and the next float up, (float) s is t below the midpoint m and
// at m itself when t's significand is even; m is exact in double.
double double_le_threshold(float t) {
  if (std::isnan(t)) return (double) t;
  if (std::isinf(t)) return t > 0 ? HUGE_VAL : -(double) FLT_MAX - std::ldexp(1.0, 103);
  double m;
  if (t == FLT_MAX) {
    m = (double) FLT_MAX + std::ldexp(1.0, 103);  // half an ulp above the top float
  } else {
    const float up = std::nextafter(t, HUGE_VALF);
    m = ((double) t + (double) up) * 0.5;
  }
  uint32_t bits;
  std::memcpy(&bits, &t, 4);
  return (bits & 1u) ? std::nextafter(m, -HUGE_VAL) : m;
}

}  // namespace abl
