/*
 * orc_near.c -- TEST INFRASTRUCTURE ONLY.  Grid helpers behind ORACLE_GRID
 * (see orc_near.h).  The cell layout is the oracle's own; the only contract
 * is that the candidate list is a superset of the agents the reference's
 * brute-force near loop accepts (src/backend/CPrinter.cpp:161-171) and is
 * visited in ascending index order.
 */
#include "orc_near.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_record_on = 0;
int orc_use_grid = 0;
long orc_record_center = 0;
static long *rec_pairs = NULL;
static size_t rec_n = 0, rec_cap = 0;

void orc_record_pair(long center, long nbr) {
  if (rec_n + 2 > rec_cap) {
    rec_cap = rec_cap ? rec_cap * 2 : 4096;
    rec_pairs = (long *) realloc(rec_pairs, rec_cap * sizeof(long));
  }
  rec_pairs[rec_n++] = center;
  rec_pairs[rec_n++] = nbr;
}

void orc_record_reset(void) { rec_n = 0; }
size_t orc_record_count(void) { return rec_n / 2; }
const long *orc_record_data(void) { return rec_pairs; }

int orc_omp_slot(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

#define ORC_MAX_THREADS 512
static orc_cands thread_cands[ORC_MAX_THREADS];

orc_cands *orc_thread_cands(int slot) { return &thread_cands[slot % ORC_MAX_THREADS]; }

static int cell_coord(double p, double lo, double cell, int n) {
  double q = floor((p - lo) / cell);
  if (!(q >= 0)) return 0; /* also catches NaN */
  if (q > (double) (n - 1)) return n - 1;
  return (int) q;
}

static size_t cell_of(const orc_grid *g, const double *p) {
  size_t c = 0;
  for (int d = g->dim - 1; d >= 0; d--) {
    c = c * (size_t) g->n[d] + (size_t) cell_coord(p[d], g->lo[d], g->cell, g->n[d]);
  }
  return c;
}

static void read_pos(const void *base, size_t i, size_t stride, size_t off, int dim,
                     int is_float, double *p) {
  const char *rec = (const char *) base + i * stride + off;
  for (int d = 0; d < dim; d++) {
    p[d] = is_float ? (double) ((const float *) rec)[d] : ((const double *) rec)[d];
  }
  for (int d = dim; d < 3; d++) p[d] = 0.0;
}

void orc_grid_build(orc_grid *g, int dim, const double lo[3], const double hi[3],
                    double granularity, double radius, const void *base,
                    size_t count, size_t stride, size_t pos_off, int pos_is_float) {
  g->dim = dim;
  /* Cells slightly wider than the granularity absorb the float rounding of
   * the reference's distance test (sqrtf of a float-rounded square sum). */
  g->cell = granularity * (1.0 + 1.0 / 1048576.0);
  g->reach = (int) ceil(radius * (1.0 + 1.0 / 1048576.0) / g->cell);
  if (g->reach < 1) g->reach = 1;
  size_t ncells = 1;
  for (int d = 0; d < 3; d++) {
    g->lo[d] = d < dim ? lo[d] : 0.0;
    int n = 1;
    if (d < dim) {
      double span = hi[d] - lo[d];
      n = (int) ceil(span / g->cell);
      if (n < 1) n = 1;
    }
    g->n[d] = n;
    ncells *= (size_t) n;
  }
  if (ncells + 1 > g->cap_cells) {
    g->cap_cells = ncells + 1;
    g->start = (int *) realloc(g->start, g->cap_cells * sizeof(int));
  }
  if (count > g->cap_items) {
    g->cap_items = count;
    g->items = (int *) realloc(g->items, g->cap_items * sizeof(int));
  }
  g->ncells = ncells;
  g->nitems = count;
  size_t *cells = (size_t *) malloc((count ? count : 1) * sizeof(size_t));
  memset(g->start, 0, (ncells + 1) * sizeof(int));
  for (size_t i = 0; i < count; i++) {
    double p[3];
    read_pos(base, i, stride, pos_off, dim, pos_is_float, p);
    cells[i] = cell_of(g, p);
    g->start[cells[i] + 1]++;
  }
  for (size_t c = 0; c < ncells; c++) g->start[c + 1] += g->start[c];
  int *fill = (int *) malloc(ncells * sizeof(int));
  memcpy(fill, g->start, ncells * sizeof(int));
  for (size_t i = 0; i < count; i++) g->items[fill[cells[i]]++] = (int) i;
  free(fill);
  free(cells);
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *) a, y = *(const int *) b;
  return (x > y) - (x < y);
}

void orc_grid_collect(const orc_grid *g, const double *p, orc_cands *out) {
  int c[3], a[3], b[3];
  for (int d = 0; d < 3; d++) {
    if (d < g->dim) {
      c[d] = cell_coord(p[d], g->lo[d], g->cell, g->n[d]);
      a[d] = c[d] - g->reach < 0 ? 0 : c[d] - g->reach;
      b[d] = c[d] + g->reach > g->n[d] - 1 ? g->n[d] - 1 : c[d] + g->reach;
    } else {
      a[d] = b[d] = 0;
    }
  }
  out->n = 0;
  for (int z = a[2]; z <= b[2]; z++) {
    for (int y = a[1]; y <= b[1]; y++) {
      size_t row = ((size_t) z * g->n[1] + (size_t) y) * g->n[0];
      int s = g->start[row + a[0]], e = g->start[row + b[0] + 1];
      size_t need = out->n + (size_t) (e - s);
      if (need > out->cap) {
        out->cap = need * 2 + 64;
        out->idx = (int *) realloc(out->idx, out->cap * sizeof(int));
      }
      for (int k = s; k < e; k++) out->idx[out->n++] = g->items[k];
    }
  }
  if (out->n <= 48) {
    for (size_t i = 1; i < out->n; i++) {
      int v = out->idx[i];
      size_t j = i;
      while (j > 0 && out->idx[j - 1] > v) { out->idx[j] = out->idx[j - 1]; j--; }
      out->idx[j] = v;
    }
  } else {
    qsort(out->idx, out->n, sizeof(int), cmp_int);
  }
}
