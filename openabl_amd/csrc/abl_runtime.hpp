// abl_runtime.hpp -- internal interface between the runtime core
// (abl_runtime.hip) and the step programs (prog_*.hip).
//
// A step program is what `-b hip` emits for one .abl model: its agent
// types, parameters, environment, and one device kernel per step function.
// The runtime owns agent storage (struct-of-arrays, two buffers per column),
// the grid pipeline and the simulate loop; a program only launches kernels on
// the column pointers the runtime hands it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "abl_device.hpp"

struct abl_sim;

namespace abl {

enum MemberKind { K_BOOL = 1, K_INT = 2, K_FLOAT = 3, K_FLOAT2 = 5, K_FLOAT3 = 6 };

inline int kind_components(int kind) {
  return kind == K_FLOAT2 ? 2 : kind == K_FLOAT3 ? 3 : 1;
}

struct MemberDesc {
  const char *name;
  int kind;
  bool is_pos;
};

struct AgentDesc {
  const char *name;
  std::vector<MemberDesc> members;
};

struct ParamDesc {
  const char *name;
  double value;  // default, as the reference prints it
};

struct StepLaunch;
struct ColSet;
struct DeltaArgs;
struct NListArgs;

// Default neighbour-list skin, as a fraction of the near radius.
constexpr double kDefaultSkin = 0.05;

// Hooks a single-type, single-step program provides so many independent runs
// of it can be batched into one launch per kernel (abl_ensemble.hip).
struct EnsembleHooks {
  size_t (*args_size)(int precision);
  // tile shape of variant `tile` (cells per tile in x, y, z)
  void (*tile_shape)(int tile, int *tx, int *ty, int *tz);
  // fill one member's kernel arguments (including its LDS capacity)
  void (*fill_args)(int precision, void *dst, const double *params, const ColSet &in,
                    const ColSet &out, const GridView &g, float near_t, float rc, int n,
                    float occupancy, int tile);
  // one launch over every member's tiles
  void (*launch_multi)(int precision, const void *d_args, const uint32_t *d_tile_base,
                       int members, unsigned total_tiles, int reach, float occupancy, int tile,
                       int block, hipStream_t stream);
  // main()'s agent records for these parameters (generator already seeded)
  int (*host_records)(const double *params, int precision, std::vector<char> &out,
                      size_t *count);
  // the gated list build (mode 1) or the step through the lists (mode 2)
  // over every member's tiles, with the members' own rebuild gates (null:
  // the program has no list form)
  void (*launch_multi_list)(int precision, const void *d_args, const uint32_t *d_tile_base,
                            int members, unsigned total_tiles, int reach, float occupancy,
                            int tile, int block, hipStream_t stream, const NListArgs &nl,
                            uint32_t *mgate, int mode);
};

struct StepDesc {
  const char *name;
  int agent;                 // stepping agent type
  int near_agent;            // agent type of the near loop, -1 if none
  const char *radius_param;  // near radius = radius_scale * param (C double math)
  double radius_scale;
  uint32_t writes;           // bit per column index of `agent` the step writes
  void (*launch)(const StepLaunch &);
  bool removes;              // calls removeCurrent()
  int adds;                  // 0, or 1 + the agent type it add()s at run time
  bool nlist;                // the launcher supports neighbour lists (StepLaunch::nl_mode)
  // near radius as the program computes it at `precision` (generated
  // programs, whose constants are folded literals); when set it replaces
  // radius_scale * radius_param
  double (*radius_fn)(int precision);
};

struct ProgramDesc {
  const char *name;
  std::vector<AgentDesc> agents;
  std::vector<ParamDesc> params;
  std::vector<StepDesc> steps;
  // environment after folding (src/AnalysisVisitor.cpp:474-541): min/max per
  // axis and the granularity, as functions of the parameters.
  void (*env)(const double *params, double lo[3], double hi[3], double *granularity);
  // main() statements before simulate(...): builds every agent type's
  // records on the host with the runtime's generator and uploads them.
  int (*host_init)(struct ::abl_sim *sim, const double *params, int precision);
  // sequential step (run on the host once per timestep after the parallel
  // steps; uses the reduction entry points), or nullptr
  int (*seq_step)(struct ::abl_sim *sim, const double *params);
  // batched-ensemble support, or nullptr
  const EnsembleHooks *ensemble;
  // has an ABL_PREC_F32C form (FP32 neighbour arithmetic, FP64 positions)
  bool f32c;
  // sequential step taking the sim's precision (generated programs), used
  // instead of seq_step when set
  int (*seq_step_prec)(struct ::abl_sim *sim, int precision);
};

// Column pointers of one agent type (members in declaration order, vector
// components expanded x, y, z).
struct ColSet {
  void *c[kMaxCols];
};

struct StepLaunch {
  hipStream_t stream;
  int precision;       // abl_precision
  int n;               // agents of the stepping type
  ColSet in;           // current values of the stepping type
  ColSet out;          // destination of written columns (others: nullptr)
  int nb_n;            // agents of the near type
  ColSet nb;           // near type, cell-sorted
  const int32_t *nb_id;
  GridView grid;
  float near_t;        // T = max{f : sqrtf(f) <= radius}
  const double *params;  // host copy, indexed as ProgramDesc::params
  int block;           // threads per block
  int order;           // 0 = cell order, 1 = reference index order
  int tile;            // LDS tile variant (0: untiled thread-per-agent kernel)
  float occupancy;     // near-type agents per cell (sizes the LDS tile budget)
  float lds_slack;     // LDS records per expected halo agent
  // neighbour lists: 0 = stencil walk, 1 = build the lists, 2 = step through
  // them, 3 = gated re-sort and build in one launch (delta, abl_pipeline.hpp)
  int nl_mode;
  const NListArgs *nl;
  const DeltaArgs *delta;
  float rc;            // near radius in cell units, widened by 2^-20
  // run-time add()/removeCurrent() and random() inside steps
  const int32_t *self_id;  // ids of the stepping agents
  uint8_t *dead;           // removeCurrent() flags, per stepping agent
  uint8_t *born;           // add() flags, per stepping agent
  ColSet birth;            // newborn columns (the added type), per parent slot
  uint64_t seed;
  uint64_t timestep;
  uint32_t step_ordinal;
};

// random() inside a step: a counter-based draw keyed by (seed, timestep,
// step ordinal, agent id, call index), so results do not depend on which
// thread runs which agent or on the order agents are stored in.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t counter_bits(uint64_t seed, uint64_t t, uint32_t step,
                                                 uint64_t id, uint32_t k) {
  uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ULL * (t + 1));
  z = mix64(z ^ (((uint64_t) step << 32) | k));
  return mix64(z + id);
}
// random_float(lo, hi) on those bits, in the build's float type
// (asset/c/libabl.c:19-24 formula).
template <typename F>
__host__ __device__ inline F counter_random(F lo, F hi, uint64_t bits) {
  const F span = (F) 18446744073709551615.0 / (hi - lo);
  return lo + (F) bits / span;
}

// Host-side helpers (abl_host.cpp).
float sqrtf_le_threshold(double r);  // max float f >= 0 with sqrtf(f) <= r
float sqrtf_lt_threshold(double r);  // min float f >= 0 with sqrtf(f) >= r
double double_le_threshold(float t);  // max double d with (float) d <= t

// Program registry (one entry per prog_*.hip).
const ProgramDesc *program_circle();
const ProgramDesc *program_circle3d();
const ProgramDesc *program_boids2d();
const ProgramDesc *program_game_of_life();
const ProgramDesc *program_predator_prey();
// registered program by name, or nullptr (abl_runtime.hip)
const ProgramDesc *find_program(const char *name);

inline double param_of(const StepLaunch &L, int idx) { return L.params[idx]; }

inline unsigned grid_blocks(int n, int block) {
  return (unsigned) ((n + block - 1) / block);
}

}  // namespace abl

#include "../../include/abl_hip.h"

namespace abl {

// The generator in the build's float type: random_float of asset/c/libabl.c
// compiled with abl_float = double, or float under use_float.
template <typename F>
inline F host_random(F lo, F hi);
template <>
inline double host_random<double>(double lo, double hi) { return abl_random_float(lo, hi); }
template <>
inline float host_random<float>(float lo, float hi) { return abl_random_floatf(lo, hi); }

}  // namespace abl
