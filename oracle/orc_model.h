/*
 * orc_model.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Every restated model under oracle/models/ describes itself with these
 * tables so one generic driver (orc_model.c) can run the C backend's simulate
 * loop (src/backend/CPrinter.cpp:189-233) and expose it to the tests.
 */
#ifndef ORACLE_ORC_MODEL_H
#define ORACLE_ORC_MODEL_H

#include "orc_near.h"

typedef struct {
  const char *name;
  char kind; /* 'i' int, 'd' abl_float, 'b' bool */
  void *ptr;
} orc_param;

typedef struct {
  const char *name;
  const type_info *info;
  dyn_array *buf;
  dyn_array *dbuf;
  size_t size;       /* sizeof(struct) */
  size_t pos_off;    /* offset of the position member */
  int dim;           /* 2 or 3, 0 when the agent has no position */
  orc_grid *grid;
} orc_agent;

typedef struct {
  const char *name;
  int agent;                 /* index into the agent table */
  void (*fn)(void *in, void *out);
  int near_agent;            /* -1 when the step has no near loop */
  double (*radius)(void);    /* value of the near radius expression */
  int removes;               /* calls removeCurrent() */
  int adds;                  /* 0, or 1 + the agent type it add()s at run time */
} orc_step;

typedef struct {
  const orc_param *params;
  const orc_agent *agents;
  int num_agents;
  const orc_step *steps;
  int num_steps;
  /* environment min/max and granularity after constant folding
   * (src/AnalysisVisitor.cpp:474-541, 1378-1399) */
  void (*env)(double lo[3], double hi[3], double *granularity);
  void (*init)(void);        /* main() statements before simulate */
  const agent_info *info;
  void *agent_struct;
  /* Models with run-time add()/removeCurrent() or a sequential step follow
   * the -b hip semantics (see orc_dynamic.c); zero for the C-backend ones. */
  int dynamic;
  void (*seq_step)(void);
} orc_model_desc;

extern const orc_model_desc orc_model;

/* ---- -b hip semantics for dynamic models (orc_dynamic.c) --------------- */
/* Within a step function: the executing agent's index and id, the timestep
 * and step ordinal (keys of the counter-based generator). */
extern __thread size_t orc_cur_index;
extern __thread long long orc_cur_id;
extern __thread unsigned orc_cur_call;
extern long long orc_timestep;
extern int orc_step_ordinal;
extern unsigned long long orc_seed;
abl_float orc_step_random(abl_float lo, abl_float hi);
void orc_remove_current(void);
void orc_add_agent(int type, const void *record);
/* Sequential-step builtins (src/main.cpp:66-130). */
int orc_count_type(int type);
long long orc_sum_member_int(int type, size_t offset, int is_bool);
void orc_log_csv_row(const char *row);
double orc_last_exec_time(void);

#endif
