// prog_boids2d.hip -- step program for examples/boids2d_cfg.abl (and, with
// steer/collision/match = 0.65/0.75/1.25 and max_speed = 1, for
// examples/boids2d.abl): the device code `-b hip` emits for update_boid.
//
// Per agent (one thread, cell-sorted order), operation for operation as the
// C backend prints it (src/backend/CPrinter.cpp:148-173,
// GenericCPrinter.cpp:20-75):
//   near loop over dist(nx.pos, in.pos) <= interaction_radius:
//     global_center += nx.pos; global_velocity += nx.velocity; count += 1
//     if (dist(in.pos, nx.pos) < separation_radius) collision_center += nx.pos, ...
//   steer/match/avoid terms, new_velocity = velocity + change * global_scale,
//   speed clamp v / |v| * max_speed, new_pos = pos + v * time_scale,
//   boundPosition wrap.
// Both distance comparisons are done on the square sum against host-computed
// thresholds, so the loop takes no square root (only |new_velocity| needs
// one).  sqrtf is monotone, so dist <= r <=> (float) s <= T and dist < r_sep
// <=> (float) s < U for float thresholds T, U; in FP64 the rounding to float
// is folded into the threshold too: (float) s <= T <=> s <= Td for the
// largest double Td that rounds to at most T.
#include <cmath>
#include <cstring>
#include <vector>

#include "abl_runtime.hpp"
#include "abl_tile.hpp"

namespace abl {
namespace {

enum { P_INTERACTION_RADIUS, P_SEPARATION_RADIUS, P_AGENT_DENSITY, P_TIME_SCALE, P_GLOBAL_SCALE,
       P_STEER_SCALE, P_COLLISION_SCALE, P_MATCH_SCALE, P_MAX_SPEED, P_NUM_AGENTS,
       P_NUM_TIMESTEPS, P_MIN_POS, P_MAX_POS };

template <typename F>
struct BoidArgs {
  const F *px, *py, *vx, *vy;  // in (cell-sorted)
  F *opx, *opy, *ovx, *ovy;    // out
  const int32_t *id;
  GridView g;
  F near_s;       // dist <= interaction_radius  <=>  s <= near_s
  F sep_s;        // dist <  separation_radius   <=>  s <= sep_s
  float rc;       // radius in cell units (row narrowing)
  int cap;        // LDS records per tile
  int n;
  F time_scale, global_scale, steer, collision, match, max_speed, min_pos, max_pos;
};

// Accumulators of the near loop.
template <typename F>
struct BoidAcc {
  F gvx = 0, gvy = 0, gcx = 0, gcy = 0, ccx = 0, ccy = 0;
  int icount = 0, ccount = 0;
  __device__ __forceinline__ void visit(const BoidArgs<F> &a, F x, F y, F qx, F qy, F qvx, F qvy) {
    // s = (nx.pos - in.pos) . (nx.pos - in.pos), as dist() forms it
    const F dx = qx - x;
    const F dy = qy - y;
    const F s = dx * dx + dy * dy;
    if (s > a.near_s) return;
    gcx = gcx + qx;
    gcy = gcy + qy;
    gvx = gvx + qvx;
    gvy = gvy + qvy;
    icount += 1;
    if (s <= a.sep_s) {
      ccx = ccx + qx;
      ccy = ccy + qy;
      ccount += 1;
    }
  }
};

// Everything after the near loop, written to slot i.
template <typename F>
__device__ __forceinline__ void boid_finish(const BoidArgs<F> &a, BoidAcc<F> s, F x, F y, F vx,
                                            F vy, uint32_t i, double *np_out = nullptr) {
  F chx = 0, chy = 0;  // velocity_change
  F sx = 0, sy = 0;    // steer_velocity
  if (s.icount > 0) {
    s.gcx = s.gcx / (F) s.icount;
    s.gcy = s.gcy / (F) s.icount;
    sx = (s.gcx - x) * a.steer;
    sy = (s.gcy - y) * a.steer;
  }
  chx = chx + sx;
  chy = chy + sy;
  F mx = 0, my = 0;    // match_velocity
  if (s.icount > 0) {
    s.gvx = s.gvx / (F) s.icount;
    s.gvy = s.gvy / (F) s.icount;
    mx = s.gvx * a.match;
    my = s.gvy * a.match;
  }
  chx = chx + mx;
  chy = chy + my;
  F ax = 0, ay = 0;    // avoid_velocity
  if (s.ccount > 0) {
    s.ccx = s.ccx / (F) s.ccount;
    s.ccy = s.ccy / (F) s.ccount;
    ax = (x - s.ccx) * a.collision;
    ay = (y - s.ccy) * a.collision;
  }
  chx = chx + ax;
  chy = chy + ay;
  F nvx = vx + chx * a.global_scale;
  F nvy = vy + chy * a.global_scale;
  const F speed = (F) sqrt_rn_f32((float) (nvx * nvx + nvy * nvy));
  if (speed > a.max_speed) {
    nvx = (nvx / speed) * a.max_speed;
    nvy = (nvy / speed) * a.max_speed;
  }
  F npx = x + nvx * a.time_scale;
  F npy = y + nvy * a.time_scale;
  npx = npx < a.min_pos ? a.max_pos : (npx > a.max_pos ? a.min_pos : npx);
  npy = npy < a.min_pos ? a.max_pos : (npy > a.max_pos ? a.min_pos : npy);
  a.opx[i] = npx;
  a.opy[i] = npy;
  a.ovx[i] = nvx;
  a.ovy[i] = nvy;
  if (np_out) {
    np_out[0] = (double) npx;
    np_out[1] = (double) npy;
  }
}

// Untiled: one thread per agent, candidates read from global memory (cell
// order), or merged into the reference's id order (BY_ID).
template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void boids_step(BoidArgs<F> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  const F x = a.px[i], y = a.py[i];
  const int3 c = cell_of<F>(a.g, x, y, (F) 0);
  BoidAcc<F> s;
  auto body = [&](uint32_t j) { s.visit(a, x, y, a.px[j], a.py[j], a.vx[j], a.vy[j]); };
  if (BY_ID) {
    for_each_candidate_by_id(a.g, c.x, c.y, 0, a.g.reach, a.id, body);
  } else {
    for_each_candidate(a.g, c.x, c.y, 0, a.g.reach, body);
  }
  boid_finish<F>(a, s, x, y, a.vx[i], a.vy[i], (uint32_t) i);
}

// LDS-tiled policy (abl_tile.hpp): records {x, y, vx, vy}; visits the same
// accepted neighbours in the same order as boids_step<F, false>.
template <typename F_>
struct BoidPolicy {
  using F = F_;
  static constexpr int DIM = 2, NS = 4;
  // register budget of the list step: 80 (six waves per SIMD) in FP64, 64
  // (eight) in FP32, as many waves as the LDS tiles of a 1M-agent population
  // leave room for; both compile without spills
  static constexpr int kListWaves = sizeof(F_) == 8 ? 6 : 8;
  using Args = BoidArgs<F>;
  __device__ static void load(const Args &a, uint32_t j, F *r) {
    r[0] = a.px[j];
    r[1] = a.py[j];
    r[2] = a.vx[j];
    r[3] = a.vy[j];
  }
  struct Acc : BoidAcc<F> {
    __device__ void visit(const Args &a, const F *own, const F *nb) {
      BoidAcc<F>::visit(a, own[0], own[1], nb[0], nb[1], nb[2], nb[3]);
    }
  };
  __device__ static void begin(const Args &, Acc &, const F *) {}
  __device__ static void finish(const Args &a, Acc &s, const F *own, uint32_t gi, double *np) {
    boid_finish<F>(a, s, own[0], own[1], own[2], own[3], gi, np);
  }
};

// Square-sum thresholds: near_t = max{f : sqrtf(f) <= interaction_radius};
// the separation test sqrtf(f) < r_sep is f < U = f <= prev(U).
template <typename F>
void set_thresholds(BoidArgs<F> &a, float near_t, double sep_radius) {
  const float u = sqrtf_lt_threshold((double) (F) sep_radius);
  const float u_le = std::nextafter(u, -HUGE_VALF);
  if (sizeof(F) == 8) {
    a.near_s = (F) double_le_threshold(near_t);
    a.sep_s = (F) double_le_threshold(u_le);
  } else {
    a.near_s = (F) near_t;
    a.sep_s = (F) u_le;
  }
}

// ---- FP32-compensated precision (ABL_PREC_F32C) -----------------------------
// Positions stay FP64 in HBM and are integrated in FP64; velocity is stored
// as float.  The near loop runs in float on offsets nx.pos - in.pos
// (|offset| <= interaction_radius): LDS records hold positions relative to
// the tile centre, and the loop accumulates the accepted offsets and
// neighbour velocities in Kahan-compensated float sums.  Since
// (sum of nx.pos) / count - pos = (sum of offsets) / count, the steer and
// avoid terms come straight from those sums.  The near test is the float
// square sum of the offset against the same thresholds, so it can differ
// from the FP64 test only for pairs within float rounding of the radius.
struct BoidArgsC {
  const double *px, *py;
  const float *vx, *vy;
  double *opx, *opy;
  float *ovx, *ovy;
  const int32_t *id;
  GridView g;
  float near_t, sep_u, rc;
  int cap, n;
  float global_scale, steer, collision, match, max_speed;
  double time_scale, min_pos, max_pos;
};

struct BoidAccC {
  KahanF dx, dy, gvx, gvy, cx, cy;
  int icount = 0, ccount = 0;
  __device__ __forceinline__ void visit(const BoidArgsC &a, float ox, float oy, float qvx,
                                        float qvy) {
    const float f = ox * ox + oy * oy;
    if (f > a.near_t) return;
    dx.add(ox);
    dy.add(oy);
    gvx.add(qvx);
    gvy.add(qvy);
    icount += 1;
    if (f < a.sep_u) {
      cx.add(ox);
      cy.add(oy);
      ccount += 1;
    }
  }
};

__device__ __forceinline__ void boid_finish_c(const BoidArgsC &a, const BoidAccC &s, double x,
                                              double y, float vx, float vy, uint32_t i,
                                              double *np_out = nullptr) {
  float chx = 0.f, chy = 0.f;
  if (s.icount > 0) {
    const float n = (float) s.icount;
    chx = chx + ((float) s.dx.value() / n) * a.steer;   // steer: mean - pos
    chy = chy + ((float) s.dy.value() / n) * a.steer;
    chx = chx + ((float) s.gvx.value() / n) * a.match;  // match: mean velocity
    chy = chy + ((float) s.gvy.value() / n) * a.match;
  }
  if (s.ccount > 0) {
    const float n = (float) s.ccount;
    chx = chx - ((float) s.cx.value() / n) * a.collision;  // avoid: pos - mean
    chy = chy - ((float) s.cy.value() / n) * a.collision;
  }
  float nvx = vx + chx * a.global_scale;
  float nvy = vy + chy * a.global_scale;
  const float speed = __builtin_sqrtf(nvx * nvx + nvy * nvy);
  if (speed > a.max_speed) {
    nvx = (nvx / speed) * a.max_speed;
    nvy = (nvy / speed) * a.max_speed;
  }
  double npx = x + (double) nvx * a.time_scale;
  double npy = y + (double) nvy * a.time_scale;
  npx = npx < a.min_pos ? a.max_pos : (npx > a.max_pos ? a.min_pos : npx);
  npy = npy < a.min_pos ? a.max_pos : (npy > a.max_pos ? a.min_pos : npy);
  a.opx[i] = npx;
  a.opy[i] = npy;
  a.ovx[i] = nvx;
  a.ovy[i] = nvy;
  if (np_out) {
    np_out[0] = npx;
    np_out[1] = npy;
  }
}

// Untiled FP32-compensated form: offsets from the FP64 difference.
template <bool BY_ID>
__global__ __launch_bounds__(1024) void boids_step_c(BoidArgsC a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  const double x = a.px[i], y = a.py[i];
  const int3 c = cell_of<double>(a.g, x, y, 0.0);
  BoidAccC s;
  auto body = [&](uint32_t j) {
    s.visit(a, (float) (a.px[j] - x), (float) (a.py[j] - y), a.vx[j], a.vy[j]);
  };
  if (BY_ID) {
    for_each_candidate_by_id(a.g, c.x, c.y, 0, a.g.reach, a.id, body);
  } else {
    for_each_candidate(a.g, c.x, c.y, 0, a.g.reach, body);
  }
  boid_finish_c(a, s, x, y, a.vx[i], a.vy[i], (uint32_t) i);
}

struct BoidPolicyC {
  using F = float;
  static constexpr int DIM = 2, NS = 4;
  static constexpr bool kRelative = true;
  using Args = BoidArgsC;
  __device__ static void load(const Args &a, uint32_t j, float *r, const double *org) {
    r[0] = (float) (a.px[j] - org[0]);
    r[1] = (float) (a.py[j] - org[1]);
    r[2] = a.vx[j];
    r[3] = a.vy[j];
  }
  __device__ static void position(const Args &a, uint32_t j, double *p) {
    p[0] = a.px[j];
    p[1] = a.py[j];
  }
  struct Acc : BoidAccC {
    float res[2];  // own record's rounding: (pos - org) - own, exact in double
    __device__ void visit(const Args &a, const float *own, const float *nb) {
      BoidAccC::visit(a, (nb[0] - own[0]) - res[0], (nb[1] - own[1]) - res[1], nb[2], nb[3]);
    }
  };
  __device__ static void begin(const Args &, Acc &s, const float *own, const double *p,
                               const double *org) {
    s.res[0] = (float) ((p[0] - org[0]) - (double) own[0]);
    s.res[1] = (float) ((p[1] - org[1]) - (double) own[1]);
  }
  __device__ static void finish(const Args &a, Acc &s, const float *own, uint32_t gi,
                                const double *p, double *np) {
    boid_finish_c(a, s, p[0], p[1], own[2], own[3], gi, np);
  }
};


void launch_c(const StepLaunch &L) {
  BoidArgsC a{};
  a.px = (const double *) L.in.c[0];
  a.py = (const double *) L.in.c[1];
  a.vx = (const float *) L.in.c[2];
  a.vy = (const float *) L.in.c[3];
  a.opx = (double *) L.out.c[0];
  a.opy = (double *) L.out.c[1];
  a.ovx = (float *) L.out.c[2];
  a.ovy = (float *) L.out.c[3];
  a.id = L.nb_id;
  a.g = L.grid;
  a.near_t = L.near_t;
  a.sep_u = sqrtf_lt_threshold(L.params[P_SEPARATION_RADIUS]);
  a.rc = L.rc;
  a.n = L.n;
  a.global_scale = (float) L.params[P_GLOBAL_SCALE];
  a.steer = (float) L.params[P_STEER_SCALE];
  a.collision = (float) L.params[P_COLLISION_SCALE];
  a.match = (float) L.params[P_MATCH_SCALE];
  a.max_speed = (float) L.params[P_MAX_SPEED];
  a.time_scale = L.params[P_TIME_SCALE];
  a.min_pos = L.params[P_MIN_POS];
  a.max_pos = L.params[P_MAX_POS];
  if (L.order == 1) {
    hipLaunchKernelGGL((boids_step_c<true>), grid_blocks(L.n, L.block), L.block, 0, L.stream, a);
  } else if (L.tile == 1 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicyC, 16, 8, 1>(L, a);
  } else if (L.tile == 2 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicyC, 8, 8, 1>(L, a);
  } else if (L.tile == 3 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicyC, 32, 8, 1>(L, a);
  } else {
    hipLaunchKernelGGL((boids_step_c<false>), grid_blocks(L.n, L.block), L.block, 0, L.stream, a);
  }
}

template <typename F>
void launch_typed(const StepLaunch &L) {
  BoidArgs<F> a{};
  a.px = (const F *) L.in.c[0];
  a.py = (const F *) L.in.c[1];
  a.vx = (const F *) L.in.c[2];
  a.vy = (const F *) L.in.c[3];
  a.opx = (F *) L.out.c[0];
  a.opy = (F *) L.out.c[1];
  a.ovx = (F *) L.out.c[2];
  a.ovy = (F *) L.out.c[3];
  a.id = L.nb_id;
  a.g = L.grid;
  set_thresholds(a, L.near_t, L.params[P_SEPARATION_RADIUS]);
  a.rc = L.rc;
  a.n = L.n;
  a.time_scale = (F) L.params[P_TIME_SCALE];
  a.global_scale = (F) L.params[P_GLOBAL_SCALE];
  a.steer = (F) L.params[P_STEER_SCALE];
  a.collision = (F) L.params[P_COLLISION_SCALE];
  a.match = (F) L.params[P_MATCH_SCALE];
  a.max_speed = (F) L.params[P_MAX_SPEED];
  a.min_pos = (F) L.params[P_MIN_POS];
  a.max_pos = (F) L.params[P_MAX_POS];
  if (L.order == 1) {
    hipLaunchKernelGGL((boids_step<F, true>), grid_blocks(L.n, L.block), L.block, 0, L.stream, a);
  } else if (L.tile == 1 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicy<F>, 16, 8, 1>(L, a);
  } else if (L.tile == 2 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicy<F>, 8, 8, 1>(L, a);
  } else if (L.tile == 3 && L.grid.reach <= 4) {
    launch_tiles<BoidPolicy<F>, 32, 8, 1>(L, a);
  } else {
    hipLaunchKernelGGL((boids_step<F, false>), grid_blocks(L.n, L.block), L.block, 0, L.stream, a);
  }
}

// main() of boids2d(_cfg).abl before simulate: num_agents times
//   add(Boid { pos: random(float2(min_pos), float2(max_pos)),
//              velocity: random(float2(-1), float2(1)) })
// members initialise in order (pos, then velocity); inside each float2 the
// y draw comes first (gcc's right-to-left argument order).
template <typename F>
void records_typed(const double *p, std::vector<char> &out, size_t *count) {
  const int n = (int) p[P_NUM_AGENTS];
  const F lo = (F) p[P_MIN_POS], hi = (F) p[P_MAX_POS];
  out.assign(4 * sizeof(F) * (size_t) (n > 0 ? n : 0), 0);
  F *aos = reinterpret_cast<F *>(out.data());
  for (int i = 0; i < n; i++) {
    const F py = host_random<F>(lo, hi);
    const F px = host_random<F>(lo, hi);
    const F vy = host_random<F>((F) (-1), (F) 1);
    const F vx = host_random<F>((F) (-1), (F) 1);
    F *r = &aos[4 * (size_t) i];
    r[0] = px;
    r[1] = py;
    r[2] = vx;
    r[3] = vy;
  }
  *count = (size_t) (n > 0 ? n : 0);
}

int host_records(const double *p, int precision, std::vector<char> &out, size_t *count) {
  if (precision == ABL_PREC_F32) records_typed<float>(p, out, count);
  else records_typed<double>(p, out, count);
  return ABL_OK;
}

int host_init(abl_sim *sim, const double *p, int precision) {
  std::vector<char> recs;
  size_t n = 0;
  host_records(p, precision, recs, &n);
  return abl_hip_upload(sim, 0, recs.data(), n);
}

// ---- batched ensembles (abl_ensemble.hip) ------------------------------------

template <typename F>
BoidArgs<F> make_args(const double *p, const ColSet &in, const ColSet &out, const GridView &g,
                      float near_t, float rc, int n) {
  BoidArgs<F> a{};
  a.px = (const F *) in.c[0];
  a.py = (const F *) in.c[1];
  a.vx = (const F *) in.c[2];
  a.vy = (const F *) in.c[3];
  a.opx = (F *) out.c[0];
  a.opy = (F *) out.c[1];
  a.ovx = (F *) out.c[2];
  a.ovy = (F *) out.c[3];
  a.g = g;
  set_thresholds(a, near_t, p[P_SEPARATION_RADIUS]);
  a.rc = rc;
  a.n = n;
  a.time_scale = (F) p[P_TIME_SCALE];
  a.global_scale = (F) p[P_GLOBAL_SCALE];
  a.steer = (F) p[P_STEER_SCALE];
  a.collision = (F) p[P_COLLISION_SCALE];
  a.match = (F) p[P_MATCH_SCALE];
  a.max_speed = (F) p[P_MAX_SPEED];
  a.min_pos = (F) p[P_MIN_POS];
  a.max_pos = (F) p[P_MAX_POS];
  return a;
}

void ens_tile_shape(int tile, int *tx, int *ty, int *tz) {
  *tz = 1;
  *ty = 8;
  *tx = tile == 2 ? 8 : tile == 3 ? 32 : 16;
}

template <typename F, int TX, int TY, int RM>
void ens_cap(float occupancy, int *cap, size_t *smem) {
  tile_budget<BoidPolicy<F>, TX, TY, 1, RM>(occupancy, cap, smem);
}

template <typename F, int TX, int TY>
void ens_cap_r(int reach, float occupancy, int *cap, size_t *smem) {
  if (reach <= 1) ens_cap<F, TX, TY, 1>(occupancy, cap, smem);
  else if (reach == 2) ens_cap<F, TX, TY, 2>(occupancy, cap, smem);
  else ens_cap<F, TX, TY, 4>(occupancy, cap, smem);
}

template <typename F>
void ens_cap_t(int tile, int reach, float occupancy, int *cap, size_t *smem) {
  if (tile == 2) ens_cap_r<F, 8, 8>(reach, occupancy, cap, smem);
  else if (tile == 3) ens_cap_r<F, 32, 8>(reach, occupancy, cap, smem);
  else ens_cap_r<F, 16, 8>(reach, occupancy, cap, smem);
}

size_t ens_args_size(int precision) {
  return precision == 0 ? sizeof(BoidArgs<double>) : sizeof(BoidArgs<float>);
}

void ens_fill_args(int precision, void *dst, const double *p, const ColSet &in, const ColSet &out,
                   const GridView &g, float near_t, float rc, int n, float occupancy, int tile) {
  size_t smem = 0;
  if (precision == 0) {
    BoidArgs<double> a = make_args<double>(p, in, out, g, near_t, rc, n);
    ens_cap_t<double>(tile, g.reach, occupancy, &a.cap, &smem);
    std::memcpy(dst, &a, sizeof a);
  } else {
    BoidArgs<float> a = make_args<float>(p, in, out, g, near_t, rc, n);
    ens_cap_t<float>(tile, g.reach, occupancy, &a.cap, &smem);
    std::memcpy(dst, &a, sizeof a);
  }
}

template <typename F, int TX, int TY, int RM>
void ens_launch(const void *args, const uint32_t *tile_base, int members, unsigned total,
                float occupancy, int block, hipStream_t st) {
  int cap = 0;
  size_t smem = 0;
  ens_cap<F, TX, TY, RM>(occupancy, &cap, &smem);
  const int threads = block > 256 ? 256 : block;
  hipLaunchKernelGGL((tiled_step_multi<BoidPolicy<F>, TX, TY, 1, RM>), total, threads, smem, st,
                     (const BoidArgs<F> *) args, tile_base, members);
}

template <typename F, int TX, int TY>
void ens_launch_r(const void *args, const uint32_t *tb, int m, unsigned total, int reach,
                  float occ, int block, hipStream_t st) {
  if (reach <= 1) ens_launch<F, TX, TY, 1>(args, tb, m, total, occ, block, st);
  else if (reach == 2) ens_launch<F, TX, TY, 2>(args, tb, m, total, occ, block, st);
  else ens_launch<F, TX, TY, 4>(args, tb, m, total, occ, block, st);
}

template <typename F>
void ens_launch_t(const void *args, const uint32_t *tb, int m, unsigned total, int reach,
                  float occ, int tile, int block, hipStream_t st) {
  if (tile == 2) ens_launch_r<F, 8, 8>(args, tb, m, total, reach, occ, block, st);
  else if (tile == 3) ens_launch_r<F, 32, 8>(args, tb, m, total, reach, occ, block, st);
  else ens_launch_r<F, 16, 8>(args, tb, m, total, reach, occ, block, st);
}

void ens_launch_multi(int precision, const void *args, const uint32_t *tb, int m, unsigned total,
                      int reach, float occ, int tile, int block, hipStream_t st) {
  if (precision == 0) ens_launch_t<double>(args, tb, m, total, reach, occ, tile, block, st);
  else ens_launch_t<float>(args, tb, m, total, reach, occ, tile, block, st);
}

template <typename F, int TX, int TY, int RM>
void ens_launch_list(const void *args, const uint32_t *tile_base, int members, unsigned total,
                     float occupancy, int block, hipStream_t st, const NListArgs &nl,
                     uint32_t *mgate, int mode) {
  int cap = 0;
  size_t smem = 0;
  ens_cap<F, TX, TY, RM>(occupancy, &cap, &smem);
  const int threads = block > 256 ? 256 : block;
  if (mode == 1)
    hipLaunchKernelGGL((tiled_list_build_multi<BoidPolicy<F>, TX, TY, 1, RM>), total, threads, smem,
                       st, (const BoidArgs<F> *) args, tile_base, members, nl, mgate);
  else
    hipLaunchKernelGGL((tiled_step_list_multi<BoidPolicy<F>, TX, TY, 1, RM>), total, threads, smem,
                       st, (const BoidArgs<F> *) args, tile_base, members, nl, mgate);
}

template <typename F, int TX, int TY>
void ens_launch_list_r(const void *args, const uint32_t *tb, int m, unsigned total, int reach,
                       float occ, int block, hipStream_t st, const NListArgs &nl, uint32_t *mg,
                       int mode) {
  if (reach <= 1) ens_launch_list<F, TX, TY, 1>(args, tb, m, total, occ, block, st, nl, mg, mode);
  else if (reach == 2) ens_launch_list<F, TX, TY, 2>(args, tb, m, total, occ, block, st, nl, mg, mode);
  else ens_launch_list<F, TX, TY, 4>(args, tb, m, total, occ, block, st, nl, mg, mode);
}

template <typename F>
void ens_launch_list_t(const void *args, const uint32_t *tb, int m, unsigned total, int reach,
                       float occ, int tile, int block, hipStream_t st, const NListArgs &nl,
                       uint32_t *mg, int mode) {
  if (tile == 2) ens_launch_list_r<F, 8, 8>(args, tb, m, total, reach, occ, block, st, nl, mg, mode);
  else if (tile == 3) ens_launch_list_r<F, 32, 8>(args, tb, m, total, reach, occ, block, st, nl, mg, mode);
  else ens_launch_list_r<F, 16, 8>(args, tb, m, total, reach, occ, block, st, nl, mg, mode);
}

void ens_launch_multi_list(int precision, const void *args, const uint32_t *tb, int m,
                           unsigned total, int reach, float occ, int tile, int block,
                           hipStream_t st, const NListArgs &nl, uint32_t *mgate, int mode) {
  if (precision == 0) ens_launch_list_t<double>(args, tb, m, total, reach, occ, tile, block, st, nl, mgate, mode);
  else ens_launch_list_t<float>(args, tb, m, total, reach, occ, tile, block, st, nl, mgate, mode);
}

const EnsembleHooks ensemble_hooks = {ens_args_size, ens_tile_shape, ens_fill_args,
                                      ens_launch_multi, host_records, ens_launch_multi_list};

void launch_update_boid(const StepLaunch &L) {
  if (L.precision == ABL_PREC_F32C) launch_c(L);
  else if (L.precision == ABL_PREC_F64) launch_typed<double>(L);
  else launch_typed<float>(L);
}

void env(const double *p, double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = p[P_MIN_POS];
  hi[0] = hi[1] = p[P_MAX_POS];
  lo[2] = hi[2] = 0.0;
  *g = p[P_INTERACTION_RADIUS];
}

}  // namespace

const ProgramDesc *program_boids2d() {
  static const ProgramDesc d = {
      "boids2d",
      {{"Boid", {{"pos", K_FLOAT2, true}, {"velocity", K_FLOAT2, false}}}},
      {{"interaction_radius", 10.0}, {"separation_radius", 2.0}, {"agent_density", 0.065536},
       {"time_scale", 0.0005}, {"global_scale", 0.15}, {"steer_scale", 1.0},
       {"collision_scale", 1.5}, {"match_scale", 1.0}, {"max_speed", 2.0},
       {"num_agents", 65536}, {"num_timesteps", 200}, {"min_pos", 0.0}, {"max_pos", 1000.0}},
      {{"update_boid", 0, 0, "interaction_radius", 1.0, 0xfu, launch_update_boid, false, 0, true}},
      env,
      host_init,
      nullptr,
      &ensemble_hooks,
      true,
  };
  return &d;
}

}  // namespace abl
