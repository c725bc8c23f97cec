// prog_circle.hip -- step program for examples/circle.abl (2-D) and
// circle3d.abl / circle3d_clustered.abl (3-D): the device code `-b hip`
// emits for step move_point.
//
// Arithmetic follows the C backend's generated expression, operation by
// operation and in the same types (src/backend/CPrinter.cpp:148-173,
// GenericCPrinter.cpp:20-75, AnalysisVisitor.cpp:89-112, 971-974):
//   new_pos = in.pos
//   for nx in near(in, 2.0*r):            dist(nx.pos, in.pos) <= 2.0*r
//     pos_dist = dist(in.pos, nx.pos)     (abl_float) sqrtf((float) sum of squares)
//     if (pos_dist == 0) continue
//     sep_dist = pos_dist - r
//     force = sep_dist > 0.0
//       ? ((nx.pos - in.pos) * (abl_float)(k_att * ((2.0*r) - pos_dist))) / pos_dist
//       : (in.pos - nx.pos) * k_rep
//     new_pos = new_pos + force
//   out.pos = clamp(new_pos, float(1), float(W - 1.0))
// The literals 2.0, 0.0 and 1.0 are C doubles, so under use_float the
// bracketed scalar expressions are evaluated in double and rounded to float
// where they are passed on, exactly as the C compiler does.
// Neighbour state is read through the cell-sorted columns; one thread per
// agent, the 3x3 (3x3x3) stencil of cells visited row by row.
#include <math.h>

#include <vector>

#include "abl_runtime.hpp"
#include "abl_tile.hpp"

namespace abl {
namespace {

enum { P_NUM_TIMESTEPS, P_NUM_AGENTS, P_RHO, P_K_REP, P_K_ATT, P_R, P_W, P_NUM_CLUSTERS,
       P_CLUSTER_SIGMA, P_BACKGROUND_FRACTION };

// Candidates of a stencil row read at once by the thread-per-agent kernels:
// the walk waits on each read in turn, so a lane's reads are issued in batches.
constexpr uint32_t kRowBatch = 4;

template <typename F, int DIM>
struct CircleArgs {
  const F *p[3];   // in.pos columns (cell-sorted)
  F *o[3];         // out.pos columns
  const int32_t *id;
  GridView g;
  float near_t;
  float rc;        // radius in cell units (row narrowing)
  int cap;         // LDS records per tile
  int n;
  F r, k_rep, k_att, W;
};

// One candidate of the near loop: new_pos += force (or skip).
template <typename F, int DIM>
__device__ __forceinline__ void circle_visit(const CircleArgs<F, DIM> &a, const F *x, const F *q,
                                             F *np) {
  // near test: dist(nx.pos, in.pos) > 2.0*r -> continue
  const float f = DIM == 3 ? dist_sq_f32<F>(q[0], q[1], q[2], x[0], x[1], x[2])
                           : dist_sq_f32<F>(q[0], q[1], x[0], x[1]);
  if (f > a.near_t) return;
  // pos_dist = dist(in.pos, nx.pos): same sum of squares (negated terms)
  const F pos_dist = (F) sqrt_rn_f32(f);
  if (pos_dist == (F) 0) return;
  const F sep_dist = pos_dist - a.r;
  F force[3];
  if ((double) sep_dist > 0.0) {
    const F s = (F) ((double) a.k_att * ((2.0 * (double) a.r) - (double) pos_dist));
    // the DIM quotients by pos_dist through one reciprocal (IEEE results)
    F num[3];
    const F y = (F) 1 / pos_dist;
    bool ok = div_ok(pos_dist);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      num[d] = (q[d] - x[d]) * s;
      force[d] = div_by_recip(num[d], pos_dist, y);
      ok = ok && div_ok(num[d]);
    }
    if (__builtin_expect(!ok, 0)) {
#pragma unroll
      for (int d = 0; d < DIM; d++) force[d] = num[d] / pos_dist;
    }
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) force[d] = (x[d] - q[d]) * a.k_rep;
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) np[d] = np[d] + force[d];
}

// out.pos = clamp(new_pos, float(1), float(W - 1.0))
template <typename F, int DIM>
__device__ __forceinline__ void circle_finish(const CircleArgs<F, DIM> &a, const F *np, uint32_t i,
                                              double *np_out = nullptr) {
  const F lo = (F) 1;
  const F hi = (F) ((double) a.W - 1.0);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const F v = np[d];
    const F c = v < lo ? lo : (v > hi ? hi : v);
    a.o[d][i] = c;
    if (np_out) np_out[d] = (double) c;
  }
}

template <typename F, int DIM, bool BY_ID>
__global__ __launch_bounds__(1024) void circle_step(CircleArgs<F, DIM> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  F x[3] = {0, 0, 0};
  F np[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    x[d] = a.p[d][i];
    np[d] = x[d];
  }
  if (BY_ID) {
    const int3 c = cell_of<F>(a.g, x[0], x[1], DIM == 3 ? x[2] : (F) 0);
    auto body = [&](uint32_t j) {
      F q[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; d++) q[d] = a.p[d][j];
      circle_visit<F, DIM>(a, x, q, np);
    };
    for_each_candidate_by_id(a.g, c.x, c.y, c.z, a.g.reach, a.id, body);
  } else {
    // stencil rows narrowed to the cells within reach; candidates read
    // through buffer resources at one 32-bit offset (a.n * sizeof(F) < 2^31,
    // checked at launch)
    ColBuf cb[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) cb[d] = col_buf(a.p[d], (uint32_t) a.n * (uint32_t) sizeof(F));
    // kRowBatch candidates of a row are read together (their loads in flight
    // at once) and then visited in order; reads past the row's end are not
    // visited, and past the column's end the buffer returns zero
    constexpr uint32_t S = (uint32_t) sizeof(F);
    for_each_candidate_row(a.g, (double) x[0], (double) x[1], DIM == 3 ? (double) x[2] : 0.0, a.rc,
                           [&](uint32_t j0, uint32_t j1) {
                             const uint32_t end = j1 * S;
                             for (uint32_t off = j0 * S; off < end; off += kRowBatch * S) {
                               F q[kRowBatch][3] = {};
#pragma unroll
                               for (int u = 0; u < kRowBatch; u++)
#pragma unroll
                                 for (int d = 0; d < DIM; d++) q[u][d] = col_load<F>(cb[d], off + u * S);
#pragma unroll
                               for (int u = 0; u < kRowBatch; u++)
                                 if (off + u * S < end) circle_visit<F, DIM>(a, x, q[u], np);
                             }
                           });
  }
  circle_finish<F, DIM>(a, np, (uint32_t) i);
}

// LDS-tiled policy (abl_tile.hpp): records are the positions.
template <typename F_, int DIM_>
struct CirclePolicy {
  using F = F_;
  static constexpr int DIM = DIM_, NS = DIM_;
  using Args = CircleArgs<F, DIM>;
  __device__ static void load(const Args &a, uint32_t j, F *r) {
#pragma unroll
    for (int d = 0; d < DIM; d++) r[d] = a.p[d][j];
  }
  struct Acc {
    F np[DIM];
    __device__ void visit(const Args &a, const F *own, const F *nb) {
      circle_visit<F, DIM>(a, own, nb, np);
    }
  };
  __device__ static void begin(const Args &, Acc &s, const F *own) {
#pragma unroll
    for (int d = 0; d < DIM; d++) s.np[d] = own[d];
  }
  __device__ static void finish(const Args &a, Acc &s, const F *, uint32_t i, double *np) {
    circle_finish<F, DIM>(a, s.np, i, np);
  }
};

// ---- FP32-compensated precision (ABL_PREC_F32C) -----------------------------
// Positions stay FP64 in HBM and the step's sum of forces is added to them in
// FP64.  The near loop runs in float on offsets nx.pos - in.pos (LDS records
// are positions relative to the tile centre): float square sum against the
// same threshold, fast float sqrt and reciprocal, and a Kahan-compensated
// float sum of the forces (up to a few hundred neighbours in dense clusters).
template <int DIM>
struct CircleArgsC {
  const double *p[3];
  double *o[3];
  const int32_t *id;
  GridView g;
  float near_t, rc;
  int cap, n;
  float r, k_rep, att_r2;  // att_r2 = 2.0 * r (folded in double, rounded)
  float k_att;
  double hi;                // W - 1.0
};

template <int DIM>
struct CircleAccC {
  KahanF f[DIM];
  // o = nx.pos - in.pos
  __device__ __forceinline__ void visit(const CircleArgsC<DIM> &a, const float *o) {
    float s2 = o[0] * o[0] + o[1] * o[1];
    if (DIM == 3) s2 = s2 + o[2] * o[2];
    if (s2 > a.near_t) return;
    if (s2 == 0.f) return;  // pos_dist == 0
    const float pos_dist = __builtin_amdgcn_sqrtf(s2);
    if (pos_dist - a.r > 0.f) {
      const float s = a.k_att * (a.att_r2 - pos_dist) * rcp_f32(pos_dist);
#pragma unroll
      for (int d = 0; d < DIM; d++) f[d].add(o[d] * s);
    } else {
#pragma unroll
      for (int d = 0; d < DIM; d++) f[d].add(-o[d] * a.k_rep);
    }
  }
};

template <int DIM>
__device__ __forceinline__ void circle_finish_c(const CircleArgsC<DIM> &a, const CircleAccC<DIM> &s,
                                                const double *x, uint32_t i,
                                                double *np_out = nullptr) {
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double v = x[d] + s.f[d].value();
    const double c = v < 1.0 ? 1.0 : (v > a.hi ? a.hi : v);
    a.o[d][i] = c;
    if (np_out) np_out[d] = c;
  }
}

template <int DIM, bool BY_ID>
__global__ __launch_bounds__(1024) void circle_step_c(CircleArgsC<DIM> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < DIM; d++) x[d] = a.p[d][i];
  CircleAccC<DIM> s;
  if (BY_ID) {
    const int3 c = cell_of<double>(a.g, x[0], x[1], x[2]);
    auto body = [&](uint32_t j) {
      float o[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < DIM; d++) o[d] = (float) (a.p[d][j] - x[d]);
      s.visit(a, o);
    };
    for_each_candidate_by_id(a.g, c.x, c.y, c.z, a.g.reach, a.id, body);
  } else {
    ColBuf cb[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) cb[d] = col_buf(a.p[d], (uint32_t) a.n * 8u);
    // batches of kRowBatch reads, as circle_step
    for_each_candidate_row(a.g, x[0], x[1], x[2], a.rc, [&](uint32_t j0, uint32_t j1) {
      const uint32_t end = j1 * 8u;
      for (uint32_t off = j0 * 8u; off < end; off += kRowBatch * 8u) {
        double q[kRowBatch][3] = {};
#pragma unroll
        for (int u = 0; u < kRowBatch; u++)
#pragma unroll
          for (int d = 0; d < DIM; d++) q[u][d] = col_load<double>(cb[d], off + u * 8u);
#pragma unroll
        for (int u = 0; u < kRowBatch; u++) {
          if (off + u * 8u >= end) break;
          float o[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int d = 0; d < DIM; d++) o[d] = (float) (q[u][d] - x[d]);
          s.visit(a, o);
        }
      }
    });
  }
  circle_finish_c<DIM>(a, s, x, (uint32_t) i);
}

template <int DIM_>
struct CirclePolicyC {
  using F = float;
  static constexpr int DIM = DIM_, NS = DIM_;
  static constexpr bool kRelative = true;
  using Args = CircleArgsC<DIM>;
  __device__ static void load(const Args &a, uint32_t j, float *r, const double *org) {
#pragma unroll
    for (int d = 0; d < DIM; d++) r[d] = (float) (a.p[d][j] - org[d]);
  }
  __device__ static void position(const Args &a, uint32_t j, double *p) {
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = a.p[d][j];
  }
  struct Acc : CircleAccC<DIM> {
    float res[DIM];  // own record's rounding: (pos - org) - own, exact in double
    __device__ void visit(const Args &a, const float *own, const float *nb) {
      float o[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < DIM; d++) o[d] = (nb[d] - own[d]) - res[d];
      CircleAccC<DIM>::visit(a, o);
    }
  };
  __device__ static void begin(const Args &, Acc &s, const float *own, const double *p,
                               const double *org) {
#pragma unroll
    for (int d = 0; d < DIM; d++) s.res[d] = (float) ((p[d] - org[d]) - (double) own[d]);
  }
  __device__ static void finish(const Args &a, Acc &s, const float *, uint32_t i, const double *p,
                                double *np) {
    circle_finish_c<DIM>(a, s, p, i, np);
  }
};


template <int DIM>
void launch_c(const StepLaunch &L) {
  CircleArgsC<DIM> a{};
  for (int d = 0; d < DIM; d++) {
    a.p[d] = (const double *) L.in.c[d];
    a.o[d] = (double *) L.out.c[d];
  }
  a.id = L.nb_id;
  a.g = L.grid;
  a.near_t = L.near_t;
  a.rc = L.rc;
  a.n = L.n;
  a.r = (float) L.params[P_R];
  a.k_rep = (float) L.params[P_K_REP];
  a.k_att = (float) L.params[P_K_ATT];
  a.att_r2 = (float) (2.0 * L.params[P_R]);
  a.hi = L.params[P_W] - 1.0;
  using M = CirclePolicyC<DIM>;
  const unsigned nb = grid_blocks(L.n, L.block);
  if (L.order == 1) {
    hipLaunchKernelGGL((circle_step_c<DIM, true>), nb, L.block, 0, L.stream, a);
  } else if (L.tile != 0 && L.grid.reach <= 4) {
    if (DIM == 3) launch_tiles<M, 8, 4, 4>(L, a);
    else if (L.tile == 2) launch_tiles<M, 8, 8, 1>(L, a);
    else if (L.tile == 3) launch_tiles<M, 32, 8, 1>(L, a);
    else launch_tiles<M, 16, 8, 1>(L, a);
  } else if ((size_t) L.n * 8u < ((size_t) 1 << 31)) {
    hipLaunchKernelGGL((circle_step_c<DIM, false>), nb, L.block, 0, L.stream, a);
  } else {
    hipLaunchKernelGGL((circle_step_c<DIM, true>), nb, L.block, 0, L.stream, a);
  }
}

template <typename F, int DIM>
void launch_typed(const StepLaunch &L) {
  CircleArgs<F, DIM> a{};
  for (int d = 0; d < DIM; d++) {
    a.p[d] = (const F *) L.in.c[d];
    a.o[d] = (F *) L.out.c[d];
  }
  a.id = L.nb_id;
  a.g = L.grid;
  a.near_t = L.near_t;
  a.rc = L.rc;
  a.n = L.n;
  a.r = (F) L.params[P_R];
  a.k_rep = (F) L.params[P_K_REP];
  a.k_att = (F) L.params[P_K_ATT];
  a.W = (F) L.params[P_W];
  const unsigned nb = grid_blocks(L.n, L.block);
  // buffer offsets of the untiled cell-order kernel are 32-bit (beyond
  // 2^31 bytes of a column the index-order kernel runs instead)
  const bool untiled_ok = (size_t) L.n * sizeof(F) < ((size_t) 1 << 31);
  if (L.order == 1) {
    hipLaunchKernelGGL((circle_step<F, DIM, true>), nb, L.block, 0, L.stream, a);
  } else if (L.tile != 0 && L.grid.reach <= 4) {
    if (DIM == 3) launch_tiles<CirclePolicy<F, DIM>, 8, 4, 4>(L, a);
    else if (L.tile == 2) launch_tiles<CirclePolicy<F, DIM>, 8, 8, 1>(L, a);
    else if (L.tile == 3) launch_tiles<CirclePolicy<F, DIM>, 32, 8, 1>(L, a);
    else launch_tiles<CirclePolicy<F, DIM>, 16, 8, 1>(L, a);
  } else if (untiled_ok) {
    hipLaunchKernelGGL((circle_step<F, DIM, false>), nb, L.block, 0, L.stream, a);
  } else {
    hipLaunchKernelGGL((circle_step<F, DIM, true>), nb, L.block, 0, L.stream, a);
  }
}

// main() of circle.abl before simulate: num_agents times
//   add(Point { pos: random(float2(W)) })
// random(float2 max) = float2(random(0, max.x), random(0, max.y)) from
// asset/lib.abl; the C backend's gcc evaluates the two calls right to left,
// so y takes the first draw and x the second.
template <typename F>
int init2(abl_sim *sim, const double *p) {
  const int n = (int) p[P_NUM_AGENTS];
  const F W = (F) p[P_W];
  std::vector<F> aos(2 * (size_t) (n > 0 ? n : 0));
  for (int i = 0; i < n; i++) {
    const F y = host_random<F>((F) 0, W);
    const F x = host_random<F>((F) 0, W);
    aos[2 * (size_t) i] = x;
    aos[2 * (size_t) i + 1] = y;
  }
  return abl_hip_upload(sim, 0, aos.data(), (size_t) n);
}

// main() of circle3d_clustered.abl before simulate (num_clusters = 0 gives
// circle3d.abl's plain random(environment.max)):
//   per_cluster = num_clusters > 0 ? int((1.0 - bg) * float(N)) / num_clusters : 0
//   for each cluster: centre = random(env.max);
//     per_cluster times: dx, dy, dz = gaussian(sigma) (in that order)
//                        add(Point { pos: clamp(centre + float3(dx,dy,dz), env.max) })
//   N - per_cluster * num_clusters times: add(Point { pos: random(env.max) })
// gaussian(s) = s * sqrt(-2 * log(1.0 - u1)) * cos(2.0 * PI * u2), u1 drawn
// first; PI is the folded literal 3.14159.  float3 random draws run z, y, x
// (gcc's right-to-left argument order).
template <typename F>
F gaussian(F sigma) {
  const F pi = (F) 3.14159;
  const F u1 = host_random<F>((F) 0, (F) 1.0);
  const F u2 = host_random<F>((F) 0, (F) 1.0);
  return (F) ((sigma * sqrt(((F) (-2)) * log((1.0 - u1)))) * cos(((2.0 * pi) * u2)));
}

template <typename F>
void random3(F W, F out[3]) {
  const F z = host_random<F>((F) 0, W);
  const F y = host_random<F>((F) 0, W);
  const F x = host_random<F>((F) 0, W);
  out[0] = x;
  out[1] = y;
  out[2] = z;
}

template <typename F>
int init3(abl_sim *sim, const double *p) {
  const int n = (int) p[P_NUM_AGENTS];
  const int nclusters = (int) p[P_NUM_CLUSTERS];
  const F W = (F) p[P_W];
  const F sigma = (F) p[P_CLUSTER_SIGMA];
  const F bg = (F) p[P_BACKGROUND_FRACTION];
  const int per_cluster =
      nclusters > 0 ? ((int) ((1.0 - bg) * ((F) n))) / nclusters : 0;
  const int nbackground = n - per_cluster * nclusters;
  std::vector<F> aos;
  aos.reserve(3 * (size_t) (n > 0 ? n : 0));
  for (int c = 0; c < nclusters; c++) {
    F centre[3];
    random3<F>(W, centre);
    for (int i = 0; i < per_cluster; i++) {
      const F dx = gaussian<F>(sigma);
      const F dy = gaussian<F>(sigma);
      const F dz = gaussian<F>(sigma);
      const F d[3] = {dx, dy, dz};
      for (int k = 0; k < 3; k++) {
        const F v = centre[k] + d[k];
        aos.push_back(v < (F) 0 ? (F) 0 : (v > W ? W : v));
      }
    }
  }
  for (int i = 0; i < nbackground; i++) {
    F q[3];
    random3<F>(W, q);
    aos.insert(aos.end(), q, q + 3);
  }
  return abl_hip_upload(sim, 0, aos.data(), aos.size() / 3);
}

int host_init2(abl_sim *sim, const double *p, int precision) {
  return precision == ABL_PREC_F32 ? init2<float>(sim, p) : init2<double>(sim, p);
}
int host_init3(abl_sim *sim, const double *p, int precision) {
  return precision == ABL_PREC_F32 ? init3<float>(sim, p) : init3<double>(sim, p);
}

template <int DIM>
void launch_move_point(const StepLaunch &L) {
  if (L.precision == ABL_PREC_F32C) launch_c<DIM>(L);
  else if (L.precision == ABL_PREC_F64) launch_typed<double, DIM>(L);
  else launch_typed<float, DIM>(L);
}

void env2(const double *p, double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = p[P_W];
  hi[2] = 0.0;
  *g = 2.0 * p[P_R];
}

void env3(const double *p, double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = hi[2] = p[P_W];
  *g = 2.0 * p[P_R];
}

}  // namespace

const ProgramDesc *program_circle() {
  static const ProgramDesc d = {
      "circle",
      {{"Point", {{"pos", K_FLOAT2, true}}}},
      {{"num_timesteps", 100}, {"num_agents", 1000}, {"rho", 0.05}, {"k_rep", 0.05},
       {"k_att", 0.01}, {"r", 5.0}, {"W", 141.421}},
      {{"move_point", 0, 0, "r", 2.0, 0x3u, launch_move_point<2>, false, 0, true}},
      env2,
      host_init2,
      nullptr,
      nullptr,
      true,
  };
  return &d;
}

const ProgramDesc *program_circle3d() {
  static const ProgramDesc d = {
      "circle3d",
      {{"Point", {{"pos", K_FLOAT3, true}}}},
      {{"num_timesteps", 100}, {"num_agents", 1000}, {"rho", 0.05}, {"k_rep", 0.05},
       {"k_att", 0.01}, {"r", 5.0}, {"W", 27.1442}, {"num_clusters", 64},
       {"cluster_sigma", 20.0}, {"background_fraction", 0.2}},
      {{"move_point", 0, 0, "r", 2.0, 0x7u, launch_move_point<3>, false, 0, true}},
      env3,
      host_init3,
      nullptr,
      nullptr,
      true,
  };
  return &d;
}

}  // namespace abl
