/*
 * orc_near.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The `for (T v : near(a, r))` loop of the C backend (src/backend/
 * CPrinter.cpp:148-173) visits every agent of type T in array-index order and
 * skips j when dist(v.pos, a.pos) > r.  The restated models write that loop as
 *
 *     ORC_NEAR_OPEN(T, v, agents.agents_T, grid_T, a->pos)
 *       if (dist_floatN(v->pos, a->pos) > r) continue;
 *       ORC_NEAR_HIT(v)
 *       { body }
 *     ORC_NEAR_CLOSE
 *
 * Without ORACLE_GRID the macros expand to exactly the reference's
 * brute-force loop.  With ORACLE_GRID the candidates come from a uniform grid
 * (cells at least as wide as the radius) and are visited in ascending array
 * index, so every inclusion test and every floating-point accumulation happens
 * in the same order as the brute-force loop: results are bit-identical, at
 * O(N*k) instead of O(N^2).
 */
#ifndef ORACLE_ORC_NEAR_H
#define ORACLE_ORC_NEAR_H

#include <stddef.h>
#include <stdint.h>

/* A counting-sorted uniform grid over one agent array. */
typedef struct {
  int dim;              /* 2 or 3 */
  double lo[3];         /* environment minimum */
  double cell;          /* cell edge */
  int n[3];             /* cells per axis */
  int reach;            /* stencil half-width in cells */
  int *start;           /* ncells + 1 offsets */
  int *items;           /* agent indices, ascending inside each cell */
  size_t nitems, cap_items;
  size_t ncells, cap_cells;
} orc_grid;

typedef struct {
  int *idx;
  size_t n, cap;
} orc_cands;

/* Build the grid for `count` records of `stride` bytes whose position member
 * (dim doubles or floats) starts at byte `pos_off`. */
void orc_grid_build(orc_grid *g, int dim, const double lo[3], const double hi[3],
                    double granularity, double radius, const void *base,
                    size_t count, size_t stride, size_t pos_off, int pos_is_float);
/* Candidates around a position, ascending by index (thread-safe). */
void orc_grid_collect(const orc_grid *g, const double *p, orc_cands *out);
orc_cands *orc_thread_cands(int slot);

/* Neighbour recording: when enabled, every agent that passes a near test is
 * appended as (center, neighbour) so tests can read the exact sets. */
extern int orc_record_on;
void orc_record_pair(long center, long nbr);
extern long orc_record_center;

#ifdef LIBABL_USE_FLOAT
#define ORC_POS_ARGS(p) double _orc_p[3] = {(double) (p).x, (double) (p).y, 0.0}
#define ORC_POS_ARGS3(p) double _orc_p[3] = {(double) (p).x, (double) (p).y, (double) (p).z}
#else
#define ORC_POS_ARGS(p) double _orc_p[3] = {(p).x, (p).y, 0.0}
#define ORC_POS_ARGS3(p) double _orc_p[3] = {(p).x, (p).y, (p).z}
#endif

#ifdef ORACLE_GRID
/* Switchable form used by the oracle library: orc_use_grid selects the grid
 * candidates; otherwise the loop is the reference's brute-force scan. */
extern int orc_use_grid;
int orc_omp_slot(void);
#define ORC_NEAR_OPEN_IMPL(T, v, arr, grid, posargs)                          \
  {                                                                           \
    orc_cands *v##_cands = NULL;                                              \
    size_t v##_count = (arr).len;                                             \
    if (orc_use_grid) {                                                       \
      posargs;                                                                \
      v##_cands = orc_thread_cands(orc_omp_slot());                           \
      orc_grid_collect(&(grid), _orc_p, v##_cands);                           \
      v##_count = v##_cands->n;                                               \
    }                                                                         \
    for (size_t v##_k = 0; v##_k < v##_count; v##_k++) {                      \
      T *v = DYN_ARRAY_GET(&(arr), T,                                         \
                           v##_cands ? (size_t) v##_cands->idx[v##_k] : v##_k);
#define ORC_NEAR_CLOSE \
  }                    \
  }
#else
/* Plain form (the reference's own loop, used for the timed CPU baseline). */
#define ORC_NEAR_OPEN_IMPL(T, v, arr, grid, posargs) \
  for (size_t v##_k = 0; v##_k < (arr).len; v##_k++) { \
    T *v = DYN_ARRAY_GET(&(arr), T, v##_k);
#define ORC_NEAR_CLOSE }
#endif

#define ORC_NEAR_OPEN(T, v, arr, grid, center_pos) \
  ORC_NEAR_OPEN_IMPL(T, v, arr, grid, ORC_POS_ARGS(center_pos))
#define ORC_NEAR_OPEN3(T, v, arr, grid, center_pos) \
  ORC_NEAR_OPEN_IMPL(T, v, arr, grid, ORC_POS_ARGS3(center_pos))

#define ORC_NEAR_HIT(v, arr, T)                                                \
  if (orc_record_on) orc_record_pair(orc_record_center,                        \
                                     (long) ((v) - (T *) (arr).values));

#endif
