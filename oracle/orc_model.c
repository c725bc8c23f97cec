/*
 * orc_model.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Generic driver compiled into every restated model.  orc_simulate() is the
 * C backend's simulate loop (src/backend/CPrinter.cpp:189-233): per timestep,
 * for each step function in order, the agent type's second buffer is copied
 * from the first exactly once (lazily, the first time it is needed), the step
 * runs for every agent i as step(&buf[i], &dbuf[i]) under an OpenMP parallel
 * for, and the two buffers are swapped.  The second buffer is never
 * re-initialised from the first after that one copy.  The sequential step is
 * not called (the C backend ignores it).
 *
 * The remaining functions are the ctypes surface the tests use.
 */
#include ABL_RT_HEADER
#include "orc_model.h"

#include <stdio.h>
#include <string.h>

extern void orc_record_reset(void);
extern size_t orc_record_count(void);
extern const long *orc_record_data(void);

#ifndef ORC_REF_BUILD
void orc_dyn_begin_step(const orc_step *st, size_t n);
void orc_dyn_set_agent(int type, size_t i);
void orc_dyn_end_step(const orc_step *st, size_t n);
void orc_ids_reset(int t, size_t n);
#endif

static const orc_param *find_param(const char *name) {
  for (const orc_param *p = orc_model.params; p->name; p++) {
    if (strcmp(p->name, name) == 0) return p;
  }
  return NULL;
}

int orc_set_param(const char *name, double value) {
  const orc_param *p = find_param(name);
  if (!p) return -1;
  switch (p->kind) {
    case 'i': *(int *) p->ptr = (int) value; break;
    case 'd': *(abl_float *) p->ptr = (abl_float) value; break;
    case 'b': *(bool *) p->ptr = value != 0.0; break;
    default: return -1;
  }
  return 0;
}

double orc_get_param(const char *name) {
  const orc_param *p = find_param(name);
  if (!p) return NAN;
  switch (p->kind) {
    case 'i': return (double) *(int *) p->ptr;
    case 'd': return (double) *(abl_float *) p->ptr;
    case 'b': return *(bool *) p->ptr ? 1.0 : 0.0;
  }
  return NAN;
}

int orc_float_bytes(void) { return (int) sizeof(abl_float); }
int orc_num_agent_types(void) { return orc_model.num_agents; }
const char *orc_agent_name(int t) { return orc_model.agents[t].name; }
size_t orc_agent_size(int t) { return orc_model.agents[t].size; }
size_t orc_agent_count(int t) { return orc_model.agents[t].buf->len; }
const void *orc_agent_data(int t) { return orc_model.agents[t].buf->values; }
int orc_num_steps(void) { return orc_model.num_steps; }
const char *orc_step_name(int s) { return orc_model.steps[s].name; }

#ifndef ORC_REF_BUILD
/* The reference's generator state is private to libabl.c; only the oracle's
 * own runtime can be reseeded. */
void orc_rng_set(uint64_t s0, uint64_t s1) {
  uint64_t s[2] = {s0, s1};
  abl_rt_rng_set(s);
}
void orc_rng_get(uint64_t *out) { abl_rt_rng_get(out); }
#endif

static void free_array(dyn_array *a) {
  free(a->values);
  a->values = NULL;
  a->len = a->cap = 0;
}

void orc_reset(void) {
  for (int t = 0; t < orc_model.num_agents; t++) {
    free_array(orc_model.agents[t].buf);
    free_array(orc_model.agents[t].dbuf);
  }
}

void orc_init(void) {
  orc_reset();
  orc_model.init();
#ifndef ORC_REF_BUILD
  for (int t = 0; t < orc_model.num_agents; t++) orc_ids_reset(t, orc_model.agents[t].buf->len);
  orc_timestep = 0;
#endif
}

/* Replace an agent array with `n` records in the generated struct layout;
 * they take ids 0..n-1. */
void orc_agent_set(int t, const void *data, size_t n) {
  const orc_agent *a = &orc_model.agents[t];
  free_array(a->buf);
  free_array(a->dbuf);
  for (size_t i = 0; i < n; i++) {
    memcpy(dyn_array_place(a->buf, a->size), (const char *) data + i * a->size, a->size);
  }
#ifndef ORC_REF_BUILD
  orc_ids_reset(t, n);
#endif
}

static void build_grid(const orc_step *st) {
  const orc_agent *na = &orc_model.agents[st->near_agent];
  double lo[3], hi[3], g;
  orc_model.env(lo, hi, &g);
  orc_grid_build(na->grid, na->dim, lo, hi, g, st->radius(), na->buf->values, na->buf->len,
                 na->size, na->pos_off, sizeof(abl_float) == sizeof(float));
}

static void run_step(int s) {
  const orc_step *st = &orc_model.steps[s];
  const orc_agent *a = &orc_model.agents[st->agent];
  if (orc_model.dynamic) {
#ifndef ORC_REF_BUILD
    /* copy-through: out starts as in (sizes may have changed) */
    free(a->dbuf->values);
    *a->dbuf = dyn_array_copy_fixed(a->size, a->buf);
    orc_step_ordinal = s;
#endif
  } else if (!a->dbuf->values) {
    *a->dbuf = dyn_array_copy_fixed(a->size, a->buf);
  }
#ifdef ORACLE_GRID
  if (orc_use_grid && st->near_agent >= 0) build_grid(st);
#endif
  char *in = (char *) a->buf->values;
  char *out = (char *) a->dbuf->values;
  size_t n = a->buf->len;
  size_t sz = a->size;
  void (*fn)(void *, void *) = st->fn;
  if (orc_model.dynamic) {
#ifndef ORC_REF_BUILD
    orc_dyn_begin_step(st, n);
    const int type = st->agent;
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) {
      orc_dyn_set_agent(type, i);
      fn(in + i * sz, out + i * sz);
    }
#endif
  } else {
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) {
      fn(in + i * sz, out + i * sz);
    }
  }
  dyn_array tmp = *a->buf;
  *a->buf = *a->dbuf;
  *a->dbuf = tmp;
#ifndef ORC_REF_BUILD
  if (orc_model.dynamic) orc_dyn_end_step(st, n);
#endif
}

void orc_simulate(int timesteps) {
  for (int t = 0; t < timesteps; t++) {
    for (int s = 0; s < orc_model.num_steps; s++) run_step(s);
#ifndef ORC_REF_BUILD
    if (orc_model.dynamic) {
      if (orc_model.seq_step) orc_model.seq_step();
      orc_timestep++;
    }
#endif
  }
}

void orc_run_step(int s) { run_step(s); }

/* Bounded sample of one step function: agents [0, m) only, each still
 * scanning the whole population (the reference loop's per-agent cost), no
 * buffer swap.  Used to time the C path where a full step would take
 * minutes. */
void orc_sample_step(int s, size_t m) {
  const orc_step *st = &orc_model.steps[s];
  const orc_agent *a = &orc_model.agents[st->agent];
  if (!a->dbuf->values) {
    *a->dbuf = dyn_array_copy_fixed(a->size, a->buf);
  }
#ifdef ORACLE_GRID
  if (orc_use_grid && st->near_agent >= 0) build_grid(st);
#endif
  char *in = (char *) a->buf->values;
  char *out = (char *) a->dbuf->values;
  size_t n = m < a->buf->len ? m : a->buf->len;
  size_t sz = a->size;
  void (*fn)(void *, void *) = st->fn;
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t i = 0; i < n; i++) {
    fn(in + i * sz, out + i * sz);
  }
}

void orc_set_grid(int on) {
#ifdef ORACLE_GRID
  orc_use_grid = on;
#else
  (void) on;
#endif
}

/* Neighbour pairs (center, neighbour) that step `s` accepts on the current
 * state, in visit order, without changing the state.  Returns the pair count;
 * copies at most `cap` pairs into `out` (2 longs each). */
long orc_neighbours(int s, long *out, long cap) {
  const orc_step *st = &orc_model.steps[s];
  const orc_agent *a = &orc_model.agents[st->agent];
  if (st->near_agent < 0) return 0;
#ifdef ORACLE_GRID
  if (orc_use_grid) build_grid(st);
#endif
  size_t n = a->buf->len;
  char *scratch = (char *) malloc(a->size * (n ? n : 1));
  if (a->dbuf->values) memcpy(scratch, a->dbuf->values, a->size * n);
  else memcpy(scratch, a->buf->values, a->size * n);
  orc_record_reset();
  orc_record_on = 1;
  for (size_t i = 0; i < n; i++) {
    orc_record_center = (long) i;
    st->fn((char *) a->buf->values + i * a->size, scratch + i * a->size);
  }
  orc_record_on = 0;
  free(scratch);
  long total = (long) orc_record_count();
  const long *d = orc_record_data();
  long m = total < cap ? total : cap;
  if (out && m > 0) memcpy(out, d, (size_t) m * 2 * sizeof(long));
  return total;
}

void orc_save(const char *path) {
  save(orc_model.agent_struct, orc_model.info, path, SAVE_JSON);
}

/* Member table of agent type t (the generated RTTI, terminated by TYPE_END
 * whose offset is sizeof(struct)). */
const type_info *orc_agent_info(int t) { return orc_model.agents[t].info; }

void orc_env(double *lo, double *hi, double *granularity) {
  orc_model.env(lo, hi, granularity);
}
double orc_step_radius(int s) {
  return orc_model.steps[s].near_agent >= 0 ? orc_model.steps[s].radius() : 0.0;
}
int orc_step_agent(int s) { return orc_model.steps[s].agent; }
int orc_step_near_agent(int s) { return orc_model.steps[s].near_agent; }
