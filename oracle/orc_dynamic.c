/*
 * orc_dynamic.c -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle's restatement of the -b hip semantics for what the C backend
 * rejects (src/backend/CBackend.cpp:30-31) or never runs
 * (src/backend/CPrinter.cpp:189-233 ignores the sequential step):
 *
 *  - out starts as a copy of in before every step (Mason's prepareOutState,
 *    src/backend/MasonPrinter.cpp:480-485), so members a step does not
 *    assign keep their values;
 *  - removeCurrent() marks the agent; dead agents stay visible to the rest of
 *    that step function and are removed, order-preserving, after it
 *    (FlameGPU's _isDead, src/backend/FlameGPUPrinter.cpp:171-173, 501-521);
 *  - add() inside a step creates at most one agent per parent
 *    (src/AnalysisVisitor.cpp:1310-1313); after the step function the
 *    newborns are appended in ascending parent id and take the next ids of
 *    their type, so they first act in later step functions (Mason schedules
 *    them for the next tick, MasonPrinter.cpp:152-174);
 *  - random() inside a step draws from a counter-based generator keyed by
 *    (seed, timestep, step ordinal, agent id, call index) instead of the
 *    shared xorshift state, so results do not depend on thread scheduling;
 *  - the sequential step runs once per timestep after the parallel steps,
 *    with count(Type), sum(Type.member), log_csv and getLastExecTime
 *    (src/main.cpp:66-130, src/backend/MasonPrinter.cpp:525-544).
 *
 * Agents keep ids (initially their creation index); the oracle's arrays stay
 * in ascending id order, which is the order its near loops visit.
 */
#include ABL_RT_HEADER
#include "orc_model.h"

#include <stdio.h>
#include <string.h>

__thread size_t orc_cur_index = 0;
__thread long long orc_cur_id = 0;
__thread unsigned orc_cur_call = 0;
long long orc_timestep = 0;
int orc_step_ordinal = 0;
unsigned long long orc_seed = 0x5EEDULL;

#define ORC_MAX_TYPES 8
static long long *ids[ORC_MAX_TYPES];
static size_t ids_cap[ORC_MAX_TYPES];
static long long next_id[ORC_MAX_TYPES];
static unsigned char *dead = NULL, *born = NULL;
static char *birth_buf = NULL;
static size_t flags_cap = 0, birth_cap = 0;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Counter-based draw: the same 64-bit value for the same key everywhere. */
uint64_t orc_counter_bits(uint64_t seed, uint64_t t, uint32_t step, uint64_t id, uint32_t k) {
  uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ULL * (t + 1));
  z = mix64(z ^ (((uint64_t) step << 32) | k));
  return mix64(z + id);
}

abl_float orc_step_random(abl_float lo, abl_float hi) {
  uint64_t bits = orc_counter_bits(orc_seed, (uint64_t) orc_timestep, (uint32_t) orc_step_ordinal,
                                   (uint64_t) orc_cur_id, orc_cur_call++);
  abl_float span = (abl_float) (UINT64_MAX / (hi - lo));
  return lo + (abl_float) bits / span;
}

void orc_remove_current(void) { dead[orc_cur_index] = 1; }

static size_t birth_size;

void orc_add_agent(int type, const void *record) {
  (void) type;
  born[orc_cur_index] = 1;
  memcpy(birth_buf + orc_cur_index * birth_size, record, birth_size);
}

void orc_ids_reset(int t, size_t n) {
  if (n > ids_cap[t]) {
    ids_cap[t] = n * 2 + 16;
    ids[t] = (long long *) realloc(ids[t], ids_cap[t] * sizeof(long long));
  }
  for (size_t i = 0; i < n; i++) ids[t][i] = (long long) i;
  next_id[t] = (long long) n;
}

const long long *orc_agent_ids(int t) { return ids[t]; }

static void ids_reserve(int t, size_t n) {
  if (n > ids_cap[t]) {
    ids_cap[t] = n * 2 + 16;
    ids[t] = (long long *) realloc(ids[t], ids_cap[t] * sizeof(long long));
  }
}

void orc_dyn_begin_step(const orc_step *st, size_t n) {
  if (n > flags_cap) {
    flags_cap = n * 2 + 16;
    dead = (unsigned char *) realloc(dead, flags_cap);
    born = (unsigned char *) realloc(born, flags_cap);
  }
  memset(dead, 0, n ? n : 1);
  memset(born, 0, n ? n : 1);
  if (st->adds) {
    birth_size = orc_model.agents[st->adds - 1].size;
    if (n * birth_size > birth_cap) {
      birth_cap = n * birth_size * 2 + 64;
      birth_buf = (char *) realloc(birth_buf, birth_cap);
    }
  }
}

void orc_dyn_set_agent(int type, size_t i) {
  orc_cur_index = i;
  orc_cur_id = ids[type][i];
  orc_cur_call = 0;
}

/* After the step function (and its buffer swap): drop the dead, keeping
 * order, then append the newborns in ascending parent index (= parent id). */
void orc_dyn_end_step(const orc_step *st, size_t n) {
  const orc_agent *a = &orc_model.agents[st->agent];
  if (st->removes) {
    size_t w = 0;
    char *base = (char *) a->buf->values;
    for (size_t i = 0; i < n; i++) {
      if (dead[i]) continue;
      if (w != i) {
        memmove(base + w * a->size, base + i * a->size, a->size);
        ids[st->agent][w] = ids[st->agent][i];
      }
      w++;
    }
    a->buf->len = w;
  }
  if (st->adds) {
    int c = st->adds - 1;
    const orc_agent *ca = &orc_model.agents[c];
    for (size_t i = 0; i < n; i++) {
      if (!born[i]) continue;
      size_t at = ca->buf->len;
      memcpy(dyn_array_place(ca->buf, ca->size), birth_buf + i * ca->size, ca->size);
      ids_reserve(c, at + 1);
      ids[c][at] = next_id[c]++;
    }
  }
}

int orc_count_type(int type) { return (int) orc_model.agents[type].buf->len; }

long long orc_sum_member_int(int type, size_t offset, int is_bool) {
  const orc_agent *a = &orc_model.agents[type];
  long long s = 0;
  for (size_t i = 0; i < a->buf->len; i++) {
    const char *p = (const char *) a->buf->values + i * a->size + offset;
    s += is_bool ? (long long) (*(const bool *) p ? 1 : 0) : (long long) *(const int *) p;
  }
  return s;
}

static char *log_text = NULL;
static size_t log_len = 0, log_cap = 0;

void orc_log_csv_row(const char *row) {
  size_t k = strlen(row);
  if (log_len + k + 1 > log_cap) {
    log_cap = (log_len + k + 1) * 2;
    log_text = (char *) realloc(log_text, log_cap);
  }
  memcpy(log_text + log_len, row, k + 1);
  log_len += k;
}

const char *orc_log_text(void) { return log_text ? log_text : ""; }
void orc_log_reset(void) { log_len = 0; if (log_text) log_text[0] = 0; }

/* The oracle does not time timesteps; getLastExecTime() reads 0. */
double orc_last_exec_time(void) { return 0.0; }

void orc_set_seed(unsigned long long s) { orc_seed = s; }
