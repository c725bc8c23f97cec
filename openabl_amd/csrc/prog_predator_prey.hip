// prog_predator_prey.hip -- step program for examples/predator_prey.abl: the
// device code `-b hip` emits for its 13 step functions and the host code of
// main()'s initialisation and of the sequential step gather_stats.
//
// Three agent types (Predator, Prey, Grass), near loops across types,
// removeCurrent() in prey_eat_or_starve / pred_eat_or_starve / prey_eaten,
// add() in pred_reproduction / prey_reproduction, random() inside steps, and a
// sequential step with count(), sum(), getLastExecTime() and log_csv().
// The C backend rejects the model (src/backend/CBackend.cpp:30-31); the
// semantics (copy-through out state, removal and births applied after each
// step function, counter-based random()) are the runtime's, documented in
// include/abl_hip.h and restated by the oracle.
//
// Every expression follows the C printer's rules (GenericCPrinter.cpp:20-75:
// scalar * vector printed as vector * scalar, unary minus as * -1.0,
// normalize = v / length(v)); distances are the reference's
// (abl_float) sqrtf((float) sum of squares) with a correctly rounded sqrt.
#include <math.h>

#include <vector>

#include "abl_runtime.hpp"

namespace abl {
namespace {

enum {
  P_SPACE_MULT, P_REPRODUCE_PREY_PROB, P_REPRODUCE_PREDATOR_PROB, P_GAIN_FROM_FOOD_PREDATOR,
  P_GAIN_FROM_FOOD_PREY, P_GRASS_REGROW_CYCLES, P_PRED_PREY_INTERACTION_RADIUS,
  P_PREY_GROUP_COHESION_RADIUS, P_SAME_SPECIES_AVOIDANCE_RADIUS, P_GRASS_EAT_DISTANCE,
  P_PRED_KILL_DISTANCE, P_DELTA_TIME, P_PRED_SPEED_ADVANTAGE, P_PREY_FRACTION,
  P_PREDATOR_FRACTION, P_AGENT_DENSITY, P_NUM_AGENTS, P_NUM_TIMESTEPS, P_ENV_SIZE
};
enum { T_PREDATOR = 0, T_PREY = 1, T_GRASS = 2 };
// columns of Predator and Prey
enum { C_PX, C_PY, C_DX, C_DY, C_SX, C_SY, C_LIFE };
// columns of Grass
enum { G_PX, G_PY, G_DEAD_CYCLES, G_AVAIL };

template <typename F>
struct Args {
  void *in[8];
  void *out[8];
  void *nb[8];
  void *birth[8];
  const int32_t *nb_id;
  const int32_t *self_id;
  uint8_t *dead;
  uint8_t *born;
  GridView g;
  float near_t;
  int n;
  uint64_t seed, t;
  uint32_t ordinal;
  F R_interact, R_cohesion, R_avoid, R_grass, R_kill, DT, PSA, SM, env, prob_prey, prob_pred;
  int gain_pred, gain_prey, regrow;
};

#define COL(set, c, T) ((T *) a.set[c])

template <typename F>
__device__ __forceinline__ F dist_of(float f) {
  return (F) sqrt_rn_f32(f);
}

// length(v) = (abl_float) sqrtf((float)(x*x + y*y))
template <typename F>
__device__ __forceinline__ F len2(F x, F y) {
  return (F) sqrt_rn_f32((float) (x * x + y * y));
}

// boundPosition(pos) = clamp(pos, float2(env_size, env_size)) = clamp(pos, 0, env)
template <typename F>
__device__ __forceinline__ F bound(F v, F env) {
  return v < (F) 0 ? (F) 0 : (v > env ? env : v);
}

template <bool BY_ID, typename Body>
__device__ __forceinline__ void near_loop(const GridView &g, const int32_t *ids, int3 c, Body &&b) {
  if (BY_ID) for_each_candidate_by_id(g, c.x, c.y, 0, g.reach, ids, b);
  else for_each_candidate(g, c.x, c.y, 0, g.reach, b);
}

#define PROLOGUE                                                   \
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);           \
  const int i = (int) (blk * blockDim.x + threadIdx.x);            \
  if (i >= a.n) return;                                            \
  const F x = COL(in, C_PX, const F)[i];                           \
  const F y = COL(in, C_PY, const F)[i];

// ---- Predator steps ------------------------------------------------------------

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void pred_follow_prey(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  F steer_x = 0, steer_y = 0;
  F cx = 0, cy = 0;
  F closest = a.R_interact;
  bool can_see = false;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F separation = dist_of<F>(f);
    if (separation < closest) {
      cx = qx[j];
      cy = qy[j];
      closest = separation;
      can_see = true;
    }
  });
  if (can_see) {
    steer_x = cx - x;
    steer_y = cy - y;
  }
  COL(out, C_SX, F)[i] = steer_x;
  COL(out, C_SY, F)[i] = steer_y;
}

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void pred_avoid(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  F ax = 0, ay = 0;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F separation = dist_of<F>(f);
    if (separation < a.R_avoid && (qx[j] != x || qy[j] != y)) {
      if ((double) separation > 0.0) {
        const F s = a.R_avoid / separation;
        ax = ax + (x - qx[j]) * s;
        ay = ay + (y - qy[j]) * s;
      }
    }
  });
  COL(out, C_SX, F)[i] = COL(in, C_SX, const F)[i] + ax;
  COL(out, C_SY, F)[i] = COL(in, C_SY, const F)[i] + ay;
}

// prey_move / pred_move: speed = the extra factor of pred_move
template <typename F, bool PRED>
__global__ __launch_bounds__(1024) void move(Args<F> a) {
  PROLOGUE
  F vx = COL(in, C_DX, const F)[i] + COL(in, C_SX, const F)[i];
  F vy = COL(in, C_DY, const F)[i] + COL(in, C_SY, const F)[i];
  const F speed = len2<F>(vx, vy);
  if ((double) speed > 1.0) {
    const F l = len2<F>(vx, vy);
    vx = vx / l;
    vy = vy / l;
  }
  F px, py;
  if (PRED) {
    px = x + ((vx * a.DT) * a.PSA) * a.SM;
    py = y + ((vy * a.DT) * a.PSA) * a.SM;
  } else {
    px = x + (vx * a.DT) * a.SM;
    py = y + (vy * a.DT) * a.SM;
  }
  COL(out, C_PX, F)[i] = bound<F>(px, a.env);
  COL(out, C_PY, F)[i] = bound<F>(py, a.env);
  COL(out, C_DX, F)[i] = vx;
  COL(out, C_DY, F)[i] = vy;
  COL(out, C_LIFE, int32_t)[i] = COL(in, C_LIFE, const int32_t)[i] - 1;
}

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void pred_eat_or_starve(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  int life = COL(in, C_LIFE, const int32_t)[i];
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    if (dist_sq_f32<F>(qx[j], qy[j], x, y) > a.near_t) return;
    life += a.gain_pred;
  });
  COL(out, C_LIFE, int32_t)[i] = life;
  if (life < 1) a.dead[i] = 1;
}

// pred_reproduction / prey_reproduction: with probability p, add a copy at
// the parent's position heading the other way and split the life.
template <typename F, bool PRED>
__global__ __launch_bounds__(1024) void reproduction(Args<F> a) {
  PROLOGUE
  const int life = COL(in, C_LIFE, const int32_t)[i];
  const F p = PRED ? a.prob_pred : a.prob_prey;
  const uint64_t bits = counter_bits(a.seed, a.t, a.ordinal, (uint64_t) a.self_id[i], 0);
  int out_life = life;
  if (counter_random<F>((F) 0, (F) 1.0, bits) < p) {
    a.born[i] = 1;
    COL(birth, C_PX, F)[i] = x;
    COL(birth, C_PY, F)[i] = y;
    COL(birth, C_DX, F)[i] = COL(in, C_DX, const F)[i] * (F) (-1.0);
    COL(birth, C_DY, F)[i] = COL(in, C_DY, const F)[i] * (F) (-1.0);
    COL(birth, C_SX, F)[i] = COL(in, C_SX, const F)[i] * (F) (-1.0);
    COL(birth, C_SY, F)[i] = COL(in, C_SY, const F)[i] * (F) (-1.0);
    COL(birth, C_LIFE, int32_t)[i] = life / 2;
    out_life = life / 2;
  }
  COL(out, C_LIFE, int32_t)[i] = out_life;
}

// ---- Prey steps ------------------------------------------------------------------

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void prey_avoid_pred(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  F ax = 0, ay = 0;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F separation = dist_of<F>(f);
    if (separation < a.R_interact) {
      if ((double) separation > 0.0) {
        const F s = a.R_interact / separation;
        ax = ax + (x - qx[j]) * s;
        ay = ay + (y - qy[j]) * s;
      }
    }
  });
  COL(out, C_SX, F)[i] = ax;
  COL(out, C_SY, F)[i] = ay;
}

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void prey_flock(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  F gcx = 0, gcy = 0, gvx = 0, gvy = 0, ax = 0, ay = 0;
  int count = 0;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F separation = dist_of<F>(f);
    gcx = gcx + qx[j];
    gcy = gcy + qy[j];
    count += 1;
    if (separation < a.R_avoid && (qx[j] != x || qy[j] != y)) {
      if ((double) separation > 0.0) {
        const F s = a.R_avoid / separation;
        ax = ax + (x - qx[j]) * s;
        ay = ay + (y - qy[j]) * s;
      }
    }
  });
  if (count > 0) {
    gcx = gcx / (F) count;
    gcy = gcy / (F) count;
    gvx = gcx - x;
    gvy = gcy - y;
  }
  COL(out, C_SX, F)[i] = (COL(in, C_SX, const F)[i] + gvx) + ax;
  COL(out, C_SY, F)[i] = (COL(in, C_SY, const F)[i] + gvy) + ay;
}

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void prey_eat_or_starve(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, G_PX, const F), *qy = COL(nb, G_PY, const F);
  const uint8_t *avail = COL(nb, G_AVAIL, const uint8_t);
  int life = COL(in, C_LIFE, const int32_t)[i];
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    if (dist_sq_f32<F>(qx[j], qy[j], x, y) > a.near_t) return;
    if (avail[j]) life += a.gain_prey;
  });
  COL(out, C_LIFE, int32_t)[i] = life;
  if (life < 1) a.dead[i] = 1;
}

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void prey_eaten(Args<F> a) {
  PROLOGUE
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  bool eaten = false;
  F closest = a.R_kill;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F distance = dist_of<F>(f);
    if (distance < closest) {
      closest = distance;
      eaten = true;
    }
  });
  if (eaten) a.dead[i] = 1;
}

// ---- Grass steps ---------------------------------------------------------------------

template <typename F, bool BY_ID>
__global__ __launch_bounds__(1024) void grass_eaten(Args<F> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  const F x = COL(in, G_PX, const F)[i];
  const F y = COL(in, G_PY, const F)[i];
  const F *qx = COL(nb, C_PX, const F), *qy = COL(nb, C_PY, const F);
  const bool in_avail = COL(in, G_AVAIL, const uint8_t)[i] != 0;
  F closest = a.R_grass;
  bool eaten = false;
  near_loop<BY_ID>(a.g, a.nb_id, cell_of<F>(a.g, x, y, (F) 0), [&](uint32_t j) {
    const float f = dist_sq_f32<F>(qx[j], qy[j], x, y);
    if (f > a.near_t) return;
    const F distance = dist_of<F>(f);
    if (distance < closest && in_avail) {
      closest = distance;
      eaten = true;
    }
  });
  int dead_cycles = COL(in, G_DEAD_CYCLES, const int32_t)[i];
  uint8_t avail = in_avail ? 1 : 0;
  if (eaten) {
    dead_cycles = 0;
    avail = 0;
  }
  COL(out, G_DEAD_CYCLES, int32_t)[i] = dead_cycles;
  COL(out, G_AVAIL, uint8_t)[i] = avail;
}

template <typename F>
__global__ __launch_bounds__(1024) void grass_growth(Args<F> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  const int in_dead = COL(in, G_DEAD_CYCLES, const int32_t)[i];
  const bool in_avail = COL(in, G_AVAIL, const uint8_t)[i] != 0;
  int dead_cycles = in_dead;
  uint8_t avail = in_avail ? 1 : 0;
  if (in_dead == a.regrow) {
    dead_cycles = 0;
    avail = 1;
  }
  if (!in_avail) dead_cycles = in_dead + 1;
  COL(out, G_DEAD_CYCLES, int32_t)[i] = dead_cycles;
  COL(out, G_AVAIL, uint8_t)[i] = avail;
}

// ---- launchers ---------------------------------------------------------------------------

template <typename F>
Args<F> make_args(const StepLaunch &L) {
  Args<F> a{};
  for (int c = 0; c < 8; c++) {
    a.in[c] = L.in.c[c];
    a.out[c] = L.out.c[c];
    a.nb[c] = L.nb.c[c];
    a.birth[c] = L.birth.c[c];
  }
  a.nb_id = L.nb_id;
  a.self_id = L.self_id;
  a.dead = L.dead;
  a.born = L.born;
  a.g = L.grid;
  a.near_t = L.near_t;
  a.n = L.n;
  a.seed = L.seed;
  a.t = L.timestep;
  a.ordinal = L.step_ordinal;
  const double *p = L.params;
  a.R_interact = (F) p[P_PRED_PREY_INTERACTION_RADIUS];
  a.R_cohesion = (F) p[P_PREY_GROUP_COHESION_RADIUS];
  a.R_avoid = (F) p[P_SAME_SPECIES_AVOIDANCE_RADIUS];
  a.R_grass = (F) p[P_GRASS_EAT_DISTANCE];
  a.R_kill = (F) p[P_PRED_KILL_DISTANCE];
  a.DT = (F) p[P_DELTA_TIME];
  a.PSA = (F) p[P_PRED_SPEED_ADVANTAGE];
  a.SM = (F) p[P_SPACE_MULT];
  a.env = (F) p[P_ENV_SIZE];
  a.prob_prey = (F) p[P_REPRODUCE_PREY_PROB];
  a.prob_pred = (F) p[P_REPRODUCE_PREDATOR_PROB];
  a.gain_pred = (int) p[P_GAIN_FROM_FOOD_PREDATOR];
  a.gain_prey = (int) p[P_GAIN_FROM_FOOD_PREY];
  a.regrow = (int) p[P_GRASS_REGROW_CYCLES];
  return a;
}

#define NEAR_LAUNCHER(name)                                                              \
  template <typename F>                                                                  \
  void launch_##name##_t(const StepLaunch &L) {                                          \
    Args<F> a = make_args<F>(L);                                                         \
    const unsigned nb = grid_blocks(L.n, L.block);                                       \
    if (L.order == 1) hipLaunchKernelGGL((name<F, true>), nb, L.block, 0, L.stream, a);  \
    else hipLaunchKernelGGL((name<F, false>), nb, L.block, 0, L.stream, a);              \
  }                                                                                      \
  void launch_##name(const StepLaunch &L) {                                              \
    if (L.precision == 0) launch_##name##_t<double>(L);                                  \
    else launch_##name##_t<float>(L);                                                    \
  }

NEAR_LAUNCHER(pred_follow_prey)
NEAR_LAUNCHER(prey_avoid_pred)
NEAR_LAUNCHER(prey_flock)
NEAR_LAUNCHER(pred_avoid)
NEAR_LAUNCHER(prey_eat_or_starve)
NEAR_LAUNCHER(pred_eat_or_starve)
NEAR_LAUNCHER(grass_eaten)
NEAR_LAUNCHER(prey_eaten)

#define PLAIN_LAUNCHER(name, kernel)                                                     \
  template <typename F>                                                                  \
  void launch_##name##_t(const StepLaunch &L) {                                          \
    Args<F> a = make_args<F>(L);                                                         \
    hipLaunchKernelGGL((kernel), grid_blocks(L.n, L.block), L.block, 0, L.stream, a);    \
  }                                                                                      \
  void launch_##name(const StepLaunch &L) {                                              \
    if (L.precision == 0) launch_##name##_t<double>(L);                                  \
    else launch_##name##_t<float>(L);                                                    \
  }

PLAIN_LAUNCHER(prey_move, (move<F, false>))
PLAIN_LAUNCHER(pred_move, (move<F, true>))
PLAIN_LAUNCHER(pred_reproduction, (reproduction<F, true>))
PLAIN_LAUNCHER(prey_reproduction, (reproduction<F, false>))
PLAIN_LAUNCHER(grass_growth, grass_growth<F>)

// ---- host: main() before simulate ----------------------------------------------------
// Grass, then Prey, then Predator, as the model's loops add them.  Member
// initialisers run in order; random(float2(env)) draws y then x (gcc's
// right-to-left arguments); randomDirection() = float2(sin(phi), cos(phi))
// with phi = random(2.0 * PI); life = GAIN + randomInt(10) = random_int(0, 10).

template <typename F>
struct Mover {
  F px, py, dx, dy, sx, sy;
  int32_t life;
};
template <typename F>
struct GrassRec {
  F px, py;
  int32_t dead_cycles;
  bool avail;
};

template <typename F>
int init_typed(abl_sim *sim, const double *p) {
  const int n = (int) p[P_NUM_AGENTS];
  const F env = (F) p[P_ENV_SIZE];
  const int num_predators = (int) ((F) p[P_PREDATOR_FRACTION] * ((F) n));
  const int num_prey = (int) ((F) p[P_PREY_FRACTION] * ((F) n));
  const int num_grass = (n - num_predators) - num_prey;
  const F pi = (F) 3.14159;
  std::vector<GrassRec<F>> grass;
  for (int i = 0; i < num_grass; i++) {
    GrassRec<F> g{};
    g.py = host_random<F>((F) 0, env);
    g.px = host_random<F>((F) 0, env);
    g.dead_cycles = 0;
    g.avail = true;
    grass.push_back(g);
  }
  auto movers = [&](int count, int gain) {
    std::vector<Mover<F>> v;
    for (int i = 0; i < count; i++) {
      Mover<F> m{};
      m.py = host_random<F>((F) 0, env);
      m.px = host_random<F>((F) 0, env);
      const F phi = host_random<F>((F) 0, (F) (2.0 * pi));
      // C's sin/cos take and return double, also under use_float.  gcc -O2
      // merges sin(phi) and cos(phi) into one sincos() call, whose result
      // can differ from separate sin/cos calls in the last bit on some
      // CPUs, so the same call is made here.
      double s_phi, c_phi;
      sincos((double) phi, &s_phi, &c_phi);
      m.dx = (F) s_phi;
      m.dy = (F) c_phi;
      m.sx = (F) 0;
      m.sy = (F) 0;
      m.life = gain + abl_random_int(0, 10);
      v.push_back(m);
    }
    return v;
  };
  std::vector<Mover<F>> prey = movers(num_prey, (int) p[P_GAIN_FROM_FOOD_PREY]);
  std::vector<Mover<F>> pred = movers(num_predators, (int) p[P_GAIN_FROM_FOOD_PREDATOR]);
  if (abl_hip_agent_size(sim, T_PREDATOR) != sizeof(Mover<F>) ||
      abl_hip_agent_size(sim, T_GRASS) != sizeof(GrassRec<F>))
    return ABL_ERR_ARG;
  int rc = abl_hip_upload(sim, T_GRASS, grass.data(), grass.size());
  if (rc) return rc;
  rc = abl_hip_upload(sim, T_PREY, prey.data(), prey.size());
  if (rc) return rc;
  return abl_hip_upload(sim, T_PREDATOR, pred.data(), pred.size());
}

int host_init(abl_sim *sim, const double *p, int precision) {
  return precision == 0 ? init_typed<double>(sim, p) : init_typed<float>(sim, p);
}

// sequential step gather_stats():
//   log_csv(count(Prey), count(Predator), sum(Grass.avail), getLastExecTime())
int gather_stats(abl_sim *sim, const double *) {
  size_t num_prey = 0, num_predator = 0;
  double avail = 0.0, exec_time = 0.0;
  int rc = abl_hip_count(sim, T_PREY, &num_prey);
  if (!rc) rc = abl_hip_count(sim, T_PREDATOR, &num_predator);
  if (!rc) rc = abl_hip_reduce_sum(sim, T_GRASS, "avail", &avail);
  if (!rc) rc = abl_hip_last_exec_time(sim, &exec_time);
  if (rc) return rc;
  char row[128];
  snprintf(row, sizeof row, "%d,%d,%d,%f\n", (int) num_prey, (int) num_predator, (int) avail,
           exec_time);
  return abl_hip_log_csv(sim, row);
}

void env(const double *p, double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = p[P_ENV_SIZE];
  hi[2] = 0.0;
  // granularity: the largest constant near radius (AnalysisVisitor.cpp:1378-1399)
  double m = p[P_PRED_PREY_INTERACTION_RADIUS];
  const int others[] = {P_PREY_GROUP_COHESION_RADIUS, P_SAME_SPECIES_AVOIDANCE_RADIUS,
                        P_GRASS_EAT_DISTANCE, P_PRED_KILL_DISTANCE};
  for (int k : others) m = p[k] > m ? p[k] : m;
  *g = m;
}

constexpr uint32_t W_STEER = (1u << C_SX) | (1u << C_SY);
constexpr uint32_t W_MOVE = (1u << C_PX) | (1u << C_PY) | (1u << C_DX) | (1u << C_DY) | (1u << C_LIFE);
constexpr uint32_t W_LIFE = 1u << C_LIFE;
constexpr uint32_t W_GRASS = (1u << G_DEAD_CYCLES) | (1u << G_AVAIL);

}  // namespace

const ProgramDesc *program_predator_prey() {
  static const AgentDesc mover_pred = {
      "Predator",
      {{"pos", K_FLOAT2, true}, {"dir", K_FLOAT2, false}, {"steer", K_FLOAT2, false},
       {"life", K_INT, false}}};
  static const AgentDesc mover_prey = {
      "Prey",
      {{"pos", K_FLOAT2, true}, {"dir", K_FLOAT2, false}, {"steer", K_FLOAT2, false},
       {"life", K_INT, false}}};
  static const AgentDesc grass = {
      "Grass", {{"pos", K_FLOAT2, true}, {"dead_cycles", K_INT, false}, {"avail", K_BOOL, false}}};
  static const ProgramDesc d = {
      "predator_prey",
      {mover_pred, mover_prey, grass},
      {{"SPACE_MULT", 1.0}, {"REPRODUCE_PREY_PROB", 0.005}, {"REPRODUCE_PREDATOR_PROB", 0.001},
       {"GAIN_FROM_FOOD_PREDATOR", 35}, {"GAIN_FROM_FOOD_PREY", 60}, {"GRASS_REGROW_CYCLES", 60},
       {"PRED_PREY_INTERACTION_RADIUS", 0.1}, {"PREY_GROUP_COHESION_RADIUS", 0.12},
       {"SAME_SPECIES_AVOIDANCE_RADIUS", 0.035}, {"GRASS_EAT_DISTANCE", 0.02},
       {"PRED_KILL_DISTANCE", 0.02}, {"DELTA_TIME", 0.0015}, {"PRED_SPEED_ADVANTAGE", 2.0},
       {"prey_fraction", 0.35}, {"predator_fraction", 0.025}, {"agent_density", 1600.0},
       {"num_agents", 3200}, {"num_timesteps", 100}, {"env_size", 1.41421}},
      {
          {"pred_follow_prey", T_PREDATOR, T_PREY, "PRED_PREY_INTERACTION_RADIUS", 1.0, W_STEER,
           launch_pred_follow_prey, false, 0},
          {"prey_avoid_pred", T_PREY, T_PREDATOR, "PRED_PREY_INTERACTION_RADIUS", 1.0, W_STEER,
           launch_prey_avoid_pred, false, 0},
          {"prey_flock", T_PREY, T_PREY, "PREY_GROUP_COHESION_RADIUS", 1.0, W_STEER,
           launch_prey_flock, false, 0},
          {"pred_avoid", T_PREDATOR, T_PREDATOR, "SAME_SPECIES_AVOIDANCE_RADIUS", 1.0, W_STEER,
           launch_pred_avoid, false, 0},
          {"prey_move", T_PREY, -1, nullptr, 0.0, W_MOVE, launch_prey_move, false, 0},
          {"pred_move", T_PREDATOR, -1, nullptr, 0.0, W_MOVE, launch_pred_move, false, 0},
          {"prey_eat_or_starve", T_PREY, T_GRASS, "GRASS_EAT_DISTANCE", 1.0, W_LIFE,
           launch_prey_eat_or_starve, true, 0},
          {"pred_eat_or_starve", T_PREDATOR, T_PREY, "PRED_KILL_DISTANCE", 1.0, W_LIFE,
           launch_pred_eat_or_starve, true, 0},
          {"grass_eaten", T_GRASS, T_PREY, "GRASS_EAT_DISTANCE", 1.0, W_GRASS,
           launch_grass_eaten, false, 0},
          {"prey_eaten", T_PREY, T_PREDATOR, "PRED_KILL_DISTANCE", 1.0, 0u, launch_prey_eaten,
           true, 0},
          {"pred_reproduction", T_PREDATOR, -1, nullptr, 0.0, W_LIFE, launch_pred_reproduction,
           false, 1 + T_PREDATOR},
          {"prey_reproduction", T_PREY, -1, nullptr, 0.0, W_LIFE, launch_prey_reproduction,
           false, 1 + T_PREY},
          {"grass_growth", T_GRASS, -1, nullptr, 0.0, W_GRASS, launch_grass_growth, false, 0},
      },
      env,
      host_init,
      gather_stats,
  };
  return &d;
}

}  // namespace abl
