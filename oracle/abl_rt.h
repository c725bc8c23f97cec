/*
 * abl_rt.h -- TEST INFRASTRUCTURE ONLY (the oracle's runtime).
 *
 * This header is the oracle's own restatement of the OpenABL C runtime
 * interface (asset/c/libabl.h:1-217).  It exposes the same names the C
 * backend's generated code calls (float2_add, dist_float2, DYN_ARRAY_GET,
 * random_float, save, ...), so the restated generated programs under
 * oracle/models/ compile unchanged against either this header or the
 * reference's own libabl.h (the oracle/_ref build).  Nothing in the product
 * (openabl_amd/) may include or link it.
 *
 * Precision follows the reference switch: abl_float is double, or float when
 * LIBABL_USE_FLOAT is defined (asset/c/libabl.h:59-63).
 */
#ifndef ORACLE_ABL_RT_H
#define ORACLE_ABL_RT_H

#include <math.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef LIBABL_USE_FLOAT
typedef float abl_float;
#else
typedef double abl_float;
#endif

/* ---- growable agent array (asset/c/libabl.h:11-57) -------------------- */

typedef struct {
  void *values;
  size_t len;
  size_t cap;
} dyn_array;

/* Capacity starts at 128 and doubles (asset/c/libabl.h:21,46-52). */
#define ABL_RT_FIRST_CAPACITY 128

static inline dyn_array dyn_array_create_fixed(size_t elem_size, size_t len) {
  dyn_array a;
  a.values = calloc(elem_size, len);
  a.len = len;
  a.cap = len;
  return a;
}

static inline dyn_array dyn_array_copy_fixed(size_t elem_size, const dyn_array *src) {
  dyn_array a = dyn_array_create_fixed(elem_size, src->len);
  if (src->len) memcpy(a.values, src->values, elem_size * src->len);
  return a;
}

static inline void *dyn_array_place(dyn_array *a, size_t elem_size) {
  if (a->len == a->cap) {
    size_t grown = a->cap == 0 ? ABL_RT_FIRST_CAPACITY : 2 * a->cap;
    a->values = realloc(a->values, elem_size * grown);
    a->cap = grown;
  }
  char *slot = (char *) a->values + elem_size * a->len;
  a->len += 1;
  return slot;
}

#define DYN_ARRAY_CREATE_FIXED(T, n) dyn_array_create_fixed(sizeof(T), (n))
#define DYN_ARRAY_COPY_FIXED(T, src) dyn_array_copy_fixed(sizeof(T), (src))
#define DYN_ARRAY_GET(a, T, i) (((T *) (a)->values) + (i))
#define DYN_ARRAY_PLACE(a, T) ((T *) dyn_array_place((a), sizeof(T)))

/* ---- scalar helpers (asset/c/libabl.h:69-74): min/max are fmin/fmax ---- */

static inline abl_float min(abl_float a, abl_float b) { return fmin(a, b); }
static inline abl_float max(abl_float a, abl_float b) { return fmax(a, b); }

/* ---- vectors (asset/c/libabl.h:80-175): component-wise, no fusion ------ */

typedef struct { abl_float x, y; } float2;
typedef struct { abl_float x, y, z; } float3;

static inline float2 float2_create(abl_float x, abl_float y) { float2 v; v.x = x; v.y = y; return v; }
static inline float2 float2_fill(abl_float s) { return float2_create(s, s); }
static inline float2 float2_add(float2 a, float2 b) { return float2_create(a.x + b.x, a.y + b.y); }
static inline float2 float2_sub(float2 a, float2 b) { return float2_create(a.x - b.x, a.y - b.y); }
static inline float2 float2_mul_scalar(float2 a, abl_float s) { return float2_create(a.x * s, a.y * s); }
static inline float2 float2_div_scalar(float2 a, abl_float s) { return float2_create(a.x / s, a.y / s); }
static inline bool float2_equals(float2 a, float2 b) { return a.x == b.x && a.y == b.y; }
static inline bool float2_not_equals(float2 a, float2 b) { return !(a.x == b.x && a.y == b.y); }

static inline float3 float3_create(abl_float x, abl_float y, abl_float z) {
  float3 v; v.x = x; v.y = y; v.z = z; return v;
}
static inline float3 float3_fill(abl_float s) { return float3_create(s, s, s); }
static inline float3 float3_add(float3 a, float3 b) { return float3_create(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline float3 float3_sub(float3 a, float3 b) { return float3_create(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float3 float3_mul_scalar(float3 a, abl_float s) { return float3_create(a.x * s, a.y * s, a.z * s); }
static inline float3 float3_div_scalar(float3 a, abl_float s) { return float3_create(a.x / s, a.y / s, a.z / s); }
static inline bool float3_equals(float3 a, float3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline bool float3_not_equals(float3 a, float3 b) { return !(a.x == b.x && a.y == b.y && a.z == b.z); }

/* Lengths go through single-precision sqrtf even when abl_float is double:
 * the double sum of squares is rounded to float, square-rooted in float and
 * widened back (asset/c/libabl.h:156-168). */
static inline abl_float dot_float2(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
static inline abl_float dot_float3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline abl_float length_float2(float2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
static inline abl_float length_float3(float3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline abl_float dist_float2(float2 a, float2 b) { return length_float2(float2_sub(a, b)); }
static inline abl_float dist_float3(float3 a, float3 b) { return length_float3(float3_sub(a, b)); }
static inline float2 normalize_float2(float2 v) { return float2_div_scalar(v, length_float2(v)); }
static inline float3 normalize_float3(float3 v) { return float3_div_scalar(v, length_float3(v)); }

/* ---- random numbers (asset/c/libabl.c:8-39) ---------------------------- */

abl_float random_float(abl_float lo, abl_float hi);
int random_int(int lo, int hi);

/* Oracle-only: read or replace the generator state (the reference has no
 * seed hook; the fixed start state is {0xdeadbeef, 0xbeefdead}). */
void abl_rt_rng_get(uint64_t state[2]);
void abl_rt_rng_set(const uint64_t state[2]);
uint64_t abl_rt_rng_next(void);

/* ---- run-time type information (asset/c/libabl.h:188-209) -------------- */

typedef enum {
  TYPE_END,
  TYPE_BOOL,
  TYPE_INT,
  TYPE_FLOAT,
  TYPE_STRING,
  TYPE_FLOAT2,
  TYPE_FLOAT3,
} type_id;

typedef struct {
  type_id type;
  unsigned offset;
  const char *name;
  bool is_pos;
} type_info;

typedef struct {
  const type_info *info;
  unsigned offset;
  const char *name;
} agent_info;

typedef enum {
  SAVE_JSON,
  SAVE_FLAME_XML,
  SAVE_FLAMEGPU_XML,
} save_type;

void save(void *agents, const agent_info *info, const char *path, save_type type);

#endif
