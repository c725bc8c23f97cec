// prog_game_of_life.hip -- step program for examples/game_of_life.abl: the
// device code `-b hip` emits for step_move.
//
//   living = 0
//   for nx in near(in, radius): living += nx.alive ? 1 : 0   (includes self)
//   alive: living -= 1; out.alive = living >= 2 && living <= 3
//   dead:  out.alive = living == 3
//   out.pos = in.pos
// The assignment out.pos = in.pos copies a member onto itself, so the
// program does not write the position columns at all (the step's write mask
// holds only `alive`): positions never change, the grid is built once, and a
// step moves 1 byte per cell plus the neighbours' flags.  Integer arithmetic:
// results are bit-exact whatever the visiting order.
#include <vector>

#include "abl_runtime.hpp"
#include "abl_tile.hpp"

namespace abl {
namespace {

enum { P_RADIUS, P_DEBUG, P_ALIVE_FRACTION, P_NUM_AGENTS, P_NUM_TIMESTEPS, P_SIZE };

template <typename F>
struct LifeArgs {
  const F *px, *py;
  const uint8_t *alive;
  uint8_t *out_alive;
  GridView g;
  float near_t;
  float rc;
  int cap;
  int n;
};

__device__ __forceinline__ uint8_t life_rule(bool alive, int living) {
  bool out;
  if (alive) {
    living -= 1;
    out = living >= 2 && living <= 3;
  } else {
    out = living == 3;
  }
  return out ? 1 : 0;
}

template <typename F>
__global__ __launch_bounds__(1024) void life_step(LifeArgs<F> a) {
  const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = (int) (blk * blockDim.x + threadIdx.x);
  if (i >= a.n) return;
  const F x = a.px[i], y = a.py[i];
  const int3 c = cell_of<F>(a.g, x, y, (F) 0);
  int living = 0;
  for_each_candidate(a.g, c.x, c.y, 0, a.g.reach, [&](uint32_t j) {
    const float f = dist_sq_f32<F>(a.px[j], a.py[j], x, y);
    if (f > a.near_t) return;
    living += a.alive[j] ? 1 : 0;
  });
  a.out_alive[i] = life_rule(a.alive[i] != 0, living);
}

// LDS-tiled policy (abl_tile.hpp): records {x, y, alive as 0/1}.
template <typename F_>
struct LifePolicy {
  using F = F_;
  static constexpr int DIM = 2, NS = 3;
  using Args = LifeArgs<F>;
  __device__ static void load(const Args &a, uint32_t j, F *r) {
    r[0] = a.px[j];
    r[1] = a.py[j];
    r[2] = a.alive[j] ? (F) 1 : (F) 0;
  }
  struct Acc {
    int living;
    __device__ void visit(const Args &a, const F *own, const F *nb) {
      if (dist_sq_f32<F>(nb[0], nb[1], own[0], own[1]) > a.near_t) return;
      living += nb[2] != (F) 0 ? 1 : 0;
    }
  };
  __device__ static void begin(const Args &, Acc &s, const F *) { s.living = 0; }
  __device__ static void finish(const Args &a, Acc &s, const F *own, uint32_t i, double *) {
    a.out_alive[i] = life_rule(own[2] != (F) 0, s.living);
  }
};

template <typename F>
void launch_typed(const StepLaunch &L) {
  LifeArgs<F> a{};
  a.px = (const F *) L.in.c[0];
  a.py = (const F *) L.in.c[1];
  a.alive = (const uint8_t *) L.in.c[2];
  a.out_alive = (uint8_t *) L.out.c[2];
  a.g = L.grid;
  a.near_t = L.near_t;
  a.rc = L.rc;
  a.n = L.n;
  using M = LifePolicy<F>;
  if (L.tile != 0 && L.grid.reach == 1) {
    launch_tiled<M, 16, 8, 1, 1>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack);
  } else if (L.tile != 0 && L.grid.reach == 2) {
    launch_tiled<M, 16, 8, 1, 2>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack);
  } else {
    hipLaunchKernelGGL(life_step<F>, grid_blocks(L.n, L.block), L.block, 0, L.stream, a);
  }
}

// main() of game_of_life.abl before simulate:
//   for x in 0..size: for y in 0..size:
//     add(Cell { pos: float2(x + 0.5, y + 0.5), alive: getInit(x, y) })
// getInit = debug ? glider-ish pattern : random(0, 1) < alive_fraction.
// Record layout: {F x, F y, bool alive} padded to the float alignment.
template <typename F>
int init_typed(abl_sim *sim, const double *p) {
  const int size = (int) p[P_SIZE];
  const bool debug = p[P_DEBUG] != 0.0;
  const F alive_fraction = (F) p[P_ALIVE_FRACTION];
  struct Rec {
    F x, y;
    bool alive;
  };
  std::vector<Rec> aos;
  aos.reserve((size_t) size * (size_t) (size > 0 ? size : 0));
  for (int x = 0; x < size; x++) {
    for (int y = 0; y < size; y++) {
      Rec r{};
      r.x = (F) (((F) x) + 0.5);
      r.y = (F) (((F) y) + 0.5);
      if (debug) {
        r.alive = (x == 10 || x == 11 || x == 12) && y == 10;
      } else {
        r.alive = host_random<F>((F) 0, (F) 1) < alive_fraction;
      }
      aos.push_back(r);
    }
  }
  if (abl_hip_agent_size(sim, 0) != sizeof(Rec)) return ABL_ERR_ARG;
  return abl_hip_upload(sim, 0, aos.data(), aos.size());
}

int host_init(abl_sim *sim, const double *p, int precision) {
  return precision == 0 ? init_typed<double>(sim, p) : init_typed<float>(sim, p);
}

void launch_step_move(const StepLaunch &L) {
  if (L.precision == 0) launch_typed<double>(L);
  else launch_typed<float>(L);
}

void env(const double *p, double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = p[P_SIZE];
  hi[2] = 0.0;
  *g = p[P_RADIUS];
}

}  // namespace

const ProgramDesc *program_game_of_life() {
  static const ProgramDesc d = {
      "game_of_life",
      {{"Cell", {{"pos", K_FLOAT2, true}, {"alive", K_BOOL, false}}}},
      {{"radius", 1.5}, {"debug", 0}, {"alive_fraction", 0.1}, {"num_agents", 400},
       {"num_timesteps", 100}, {"size", 20}},
      {{"step_move", 0, 0, "radius", 1.0, 0x4u, launch_step_move}},
      env,
      host_init,
  };
  return &d;
}

}  // namespace abl
