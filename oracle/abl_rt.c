/*
 * abl_rt.c -- TEST INFRASTRUCTURE ONLY (the oracle's runtime).
 *
 * Restates the OpenABL C runtime's generator and JSON writer:
 *   - xorshift128+ with the fixed start state {0xdeadbeef, 0xbeefdead}
 *     (asset/c/libabl.c:8-17),
 *   - random_float / random_int (asset/c/libabl.c:19-39), including the
 *     power-of-two branch of random_int that returns x & (n-1) without
 *     adding the lower bound,
 *   - save() in JSON form (asset/c/libabl.c:46-124, 215-231): members with
 *     "%f", vectors as [x,y(,z)], bools as true/false, agents separated by
 *     ",\n", agent types and members in declaration order.
 * Only tests, the smoke check and the benchmark's CPU leg use this file.
 */
#include "abl_rt.h"

#include <limits.h>
#include <stdio.h>

static uint64_t rng_s[2] = {0xdeadbeefULL, 0xbeefdeadULL};

uint64_t abl_rt_rng_next(void) {
  uint64_t a = rng_s[0];
  const uint64_t b = rng_s[1];
  rng_s[0] = b;
  a ^= a << 23;
  rng_s[1] = a ^ b ^ (a >> 17) ^ (b >> 26);
  return rng_s[1] + b;
}

void abl_rt_rng_get(uint64_t state[2]) { state[0] = rng_s[0]; state[1] = rng_s[1]; }
void abl_rt_rng_set(const uint64_t state[2]) { rng_s[0] = state[0]; rng_s[1] = state[1]; }

abl_float random_float(abl_float lo, abl_float hi) {
  uint64_t bits = abl_rt_rng_next();
  /* UINT64_MAX converts to 2^64 in either precision; the quotient is taken
   * in abl_float exactly as the reference writes it. */
  abl_float span = (abl_float) (UINT64_MAX / (hi - lo));
  return lo + (abl_float) bits / span;
}

int random_int(int lo, int hi) {
  unsigned n = (unsigned) (hi - lo + 1);
  if ((n & (n - 1)) == 0) {
    return (int) (abl_rt_rng_next() & (n - 1));
  }
  unsigned reject_from = UINT_MAX - UINT_MAX % n;
  unsigned draw;
  do {
    draw = (unsigned) abl_rt_rng_next();
  } while (draw >= reject_from);
  return lo + (int) (draw % n);
}

/* ---- JSON writer -------------------------------------------------------- */

static size_t record_size(const type_info *m) {
  while (m->type != TYPE_END) m++;
  return m->offset; /* the terminator carries sizeof(struct) */
}

static void put_member(FILE *f, const char *rec, const type_info *m) {
  const char *p = rec + m->offset;
  switch (m->type) {
    case TYPE_BOOL: fputs(*(const bool *) p ? "true" : "false", f); break;
    case TYPE_INT: fprintf(f, "%d", *(const int *) p); break;
    case TYPE_FLOAT: fprintf(f, "%f", (double) *(const abl_float *) p); break;
    case TYPE_FLOAT2: {
      const float2 *v = (const float2 *) p;
      fprintf(f, "[%f,%f]", (double) v->x, (double) v->y);
      break;
    }
    case TYPE_FLOAT3: {
      const float3 *v = (const float3 *) p;
      fprintf(f, "[%f,%f,%f]", (double) v->x, (double) v->y, (double) v->z);
      break;
    }
    default: abort();
  }
}

static void put_records(FILE *f, const dyn_array *arr, const type_info *members) {
  size_t sz = record_size(members);
  fputc('[', f);
  for (size_t i = 0; i < arr->len; i++) {
    if (i) fputs(",\n", f);
    const char *rec = (const char *) arr->values + sz * i;
    fputc('{', f);
    for (const type_info *m = members; m->type != TYPE_END; m++) {
      if (m != members) fputc(',', f);
      fprintf(f, "\"%s\":", m->name);
      put_member(f, rec, m);
    }
    fputc('}', f);
  }
  fputc(']', f);
}

void save(void *agents, const agent_info *info, const char *path, save_type type) {
  FILE *f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "save(): Count not open \"%s\" for writing\n", path);
    return;
  }
  if (type == SAVE_JSON) {
    fputc('{', f);
    for (const agent_info *a = info; a->name; a++) {
      if (a != info) fputc(',', f);
      fprintf(f, "\"%s\":", a->name);
      put_records(f, (const dyn_array *) ((const char *) agents + a->offset), a->info);
    }
    fputc('}', f);
  }
  fclose(f);
}
