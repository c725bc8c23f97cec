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
} orc_model_desc;

extern const orc_model_desc orc_model;

#endif
