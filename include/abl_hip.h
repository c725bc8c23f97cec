/*
 * abl_hip.h -- C ABI of the MI355X (gfx950) OpenABL runtime.
 *
 * This is the run-time boundary between code generated for `-b hip` and the
 * device runtime.  It replaces, for that backend, the reference's run-time
 * interface asset/c/libabl.h:1-217 (generator, RTTI, save) and the simulate
 * loop the C backend prints inline (src/backend/CPrinter.cpp:148-233).
 * Every entry point is plain C: opaque handles, pointers and sizes, an int
 * status (0 = ABL_OK) and a thread-local message from abl_hip_last_error().
 *
 * Agent data crosses the boundary in exactly the C backend's struct layout
 * (src/backend/CPrinter.cpp:250-265: members in declaration order, natural
 * alignment, float members as double or, with use_float, float), so host
 * initialisation code and save() stay byte-identical to the C backend.
 * Inside the runtime agents live in HBM as struct-of-arrays, double-buffered,
 * sorted by grid cell.
 *
 * Thread safety: distinct handles may be used from distinct threads
 * concurrently; one handle must not be used from two threads at once.
 */
#ifndef ABL_HIP_H
#define ABL_HIP_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  ABL_OK = 0,
  ABL_ERR_ARG = 1,          /* invalid argument (unknown name, bad index, size) */
  ABL_ERR_STATE = 2,        /* call not valid in the handle's current state */
  ABL_ERR_DEVICE = 3,       /* HIP runtime error (message names the call) */
  ABL_ERR_MEMORY = 4,       /* device or host allocation failed / over budget */
  ABL_ERR_UNSUPPORTED = 5,  /* feature not supported (the reference's exit 2,
                               src/main.cpp:270-274) */
};

/* Human-readable message of the last failing call on this thread. */
const char *abl_hip_last_error(void);

/* ---- run-time type information ----------------------------------------
 * Same values and layout as the reference's type_id / type_info /
 * agent_info (asset/c/libabl.h:188-209), so generated RTTI tables can be
 * passed unchanged. */
typedef enum {
  ABL_TYPE_END = 0,
  ABL_TYPE_BOOL = 1,
  ABL_TYPE_INT = 2,
  ABL_TYPE_FLOAT = 3,
  ABL_TYPE_STRING = 4,
  ABL_TYPE_FLOAT2 = 5,
  ABL_TYPE_FLOAT3 = 6,
} abl_type_id;

typedef struct {
  abl_type_id type;
  unsigned offset;
  const char *name;
  bool is_pos;
} abl_type_info;

typedef struct {
  const abl_type_info *info;
  unsigned offset;
  const char *name;
} abl_agent_info;

/* The reference's dyn_array (asset/c/libabl.h:11-15). */
typedef struct {
  void *values;
  size_t len;
  size_t cap;
} abl_dyn_array;

/* ---- host runtime (replaces asset/c/libabl.c for -b hip) ---------------
 * xorshift128+ from the fixed state {0xdeadbeef, 0xbeefdead}
 * (asset/c/libabl.c:8-39).  The generator is per process, like the
 * reference's; abl_rng_seed() replaces the state (the `hip.seed` config). */
double abl_random_float(double lo, double hi);      /* FP64 build semantics */
float abl_random_floatf(float lo, float hi);         /* use_float semantics */
int abl_random_int(int lo, int hi);
void abl_rng_seed(uint64_t s0, uint64_t s1);
void abl_rng_state(uint64_t out[2]);

/* save(agents, info, path, SAVE_JSON) (asset/c/libabl.c:46-124, 215-231):
 * `agents` points at the generated `struct agent_struct` whose dyn_arrays
 * the agent_info offsets address; float_bytes is 8, or 4 under use_float.
 * Returns ABL_ERR_ARG if the file cannot be opened (the reference prints to
 * stderr and returns). */
int abl_save_json(const void *agents, const abl_agent_info *info, const char *path,
                  int float_bytes);

/* save(agents, info, path, SAVE_FLAME_XML) (for_gpu = 0) or SAVE_FLAMEGPU_XML
 * (for_gpu = 1) (asset/c/libabl.c:126-213): the FLAME / FlameGPU state file
 * of the agents, arguments as for abl_save_json.  ABL_ERR_ARG also when
 * for_gpu and a position member is not a vector. */
int abl_save_flame_xml(const void *agents, const abl_agent_info *info, const char *path,
                       int float_bytes, int for_gpu);

/* ---- simulations ------------------------------------------------------- */
typedef struct abl_sim abl_sim;

typedef enum {
  ABL_PREC_F64 = 0,   /* the C backend's default (double) */
  ABL_PREC_F32 = 1,   /* -C use_float=true (asset/c/libabl.h:59-63) */
  ABL_PREC_F32C = 2,  /* FP32 neighbour arithmetic with compensated sums,
                         FP64 positions and integration */
} abl_precision;

/* Step programs compiled into this library (the device code `-b hip` emits
 * per model).  Names: "circle", "circle3d", "boids2d", "game_of_life". */
int abl_hip_program_count(void);
const char *abl_hip_program_name(int index);

/* Create a simulation of `program` at `precision` on the current device.
 * Parameters start at the model's defaults. */
int abl_hip_create(const char *program, int precision, abl_sim **out);
int abl_hip_destroy(abl_sim *sim);

/* `-P name=value`: set a model parameter (also derived constants, which the
 * caller folds as the reference compiler does).  Only before the first
 * simulate call after an upload. */
int abl_hip_set_param(abl_sim *sim, const char *name, double value);
int abl_hip_get_param(abl_sim *sim, const char *name, double *value);

/* `-C key=value` runtime options: "hip.block" (threads per block),
 * "hip.cell_scale" (cells per granularity, 1..4), "hip.order" ("cell" or
 * "index": visit neighbours in cell order or in the reference's index
 * order), "hip.tile" (LDS tile shape, 0 = untiled), "hip.lds_slack" (LDS
 * records per tile as a multiple of the expected halo), "hip.grid" ("delta"
 * or "full" re-sort), "hip.nlist" (1: neighbour lists with a skin for tiled
 * near loops over the agent's own type, 0: walk the stencil every step;
 * left unset, lists are used and given up for a type whose lists were
 * rebuilt on more than half of the 16 timesteps after their first build --
 * decided at that fixed timestep, so results do not depend on timing),
 * "hip.skin" (list skin as a fraction of the radius), "hip.list_cap" (list
 * entries per agent; a longer list falls back to the stencil walk),
 * "hip.cell_inflate" (cell widening factor; list mode widens by 1 + skin),
 * "hip.rebuild" ("fused", the default: a list rebuild's re-sort and build
 * run as one launch whose phases are ordered by a work queue -- blocks claim
 * items in order and wait only for items already claimed by running blocks,
 * so every wait ends whatever the launch size and whatever else runs on the
 * device; "split": four launches; the two are bit-identical),
 * "hip.rebuild_blocks" (blocks of that launch, 0 = as many as the CUs hold at
 * once; larger values only add blocks that find the queue empty),
 * "hip.barrier_spins" (spin limit of one phase wait, far above any phase; a
 * limit forced below a phase's duration abandons the rebuild, the next
 * synchronize or download returns ABL_ERR_DEVICE and the runtime rebuilds
 * from scratch with the split launches from then on),
 * "hip.timing" (1: time kernels with HIP events), "hip.seed" (key of
 * the counter-based generator random() uses inside steps), "hip.log_csv"
 * (file log_csv rows are appended to), "hip.defer_upload" ("1": from now on
 * abl_hip_upload and abl_hip_init_agents keep the records on the host and
 * make no device call, so another thread can build the initial state; "0":
 * upload what was kept and return to direct uploads). */
int abl_hip_set_config(abl_sim *sim, const char *key, const char *value);

/* Agent types of the program, in declaration order. */
int abl_hip_agent_type_count(abl_sim *sim);
const char *abl_hip_agent_type_name(abl_sim *sim, int agent_type);
/* Generated-struct size of an agent type at the sim's precision. */
size_t abl_hip_agent_size(abl_sim *sim, int agent_type);

/* Replace the agents of one type with `count` records in the generated struct
 * layout (the host's `agents.agents_T.values`).  Record i becomes agent id i. */
int abl_hip_upload(abl_sim *sim, int agent_type, const void *aos, size_t count);
/* Copy the agents of one type back in id order (the reference's array
 * order).  `capacity` is in records; `count` receives the number alive. */
int abl_hip_download(abl_sim *sim, int agent_type, void *aos, size_t capacity,
                     size_t *count);
int abl_hip_count(abl_sim *sim, int agent_type, size_t *count);

/* Run the program's main() statements before simulate(...) on the host --
 * the agent-creation loops, drawing from the process generator
 * (abl_random_float) in the C backend's evaluation order -- and upload the
 * result.  Replaces the init half of the generated main()
 * (src/backend/CPrinter.cpp:69-73, 84-90, 267-272). */
int abl_hip_init_agents(abl_sim *sim);

/* Upload / download every type through the generated agent_struct and
 * agent_info table (the same arguments save() takes). */
int abl_hip_upload_all(abl_sim *sim, const void *agents, const abl_agent_info *info);
int abl_hip_download_all(abl_sim *sim, void *agents, const abl_agent_info *info);

/* Run `timesteps` timesteps of every step function in program order
 * (src/backend/CPrinter.cpp:189-233).  Asynchronous on the sim's stream. */
int abl_hip_simulate(abl_sim *sim, int timesteps);
/* Wait for the sim's stream. */
int abl_hip_synchronize(abl_sim *sim);
/* Non-blocking: *done = 1 when all work queued on the sim has finished. */
int abl_hip_query(abl_sim *sim, int *done);
/* An upper bound of the device bytes a simulation of `program` with these
 * parameters (-P overrides; the rest at their defaults) would hold at any
 * time while it has at most `agents` agents of each type: columns, ids, grid
 * and sort scratch (at least `cells` cells, and the grid of the finest cell
 * size), neighbour lists at the largest capacity the runtime grows them to,
 * add()/removeCurrent() flags and newborn staging, upload/download staging.
 * Admission control for serving; abl_hip_footprint() of a live simulation
 * never exceeds it. */
int abl_hip_estimate_footprint(const char *program, int precision, int num_params,
                               const char *const *param_names, const double *param_values,
                               size_t agents, size_t cells, size_t *bytes);

/* Parity probe: the (centre id, neighbour id) pairs that step function
 * `step`'s near loop accepts on the current state, sorted by centre then
 * neighbour id.  Pass pairs == NULL to query the count. */
int abl_hip_neighbours(abl_sim *sim, int step, int64_t *pairs, size_t capacity,
                       size_t *count);

/* Ids of one type's agents, ascending (the order download() returns them).
 * Ids start as the upload index; after run-time add()/removeCurrent() they
 * are sparse: survivors keep theirs and newborns take the next ids of their
 * type in ascending parent id. */
int abl_hip_download_ids(abl_sim *sim, int agent_type, int64_t *ids, size_t capacity,
                         size_t *count);

/* ---- sequential-step builtins (src/main.cpp:66-130) -------------------
 * count(Type) is abl_hip_count.  sum(Type.member) sums per component (bools
 * count true values) into out[0..components); integer sums are exact and
 * floating sums are added in a fixed order, so they are deterministic.
 * count(Type.member, value) counts scalar members equal to value.
 * getLastExecTime() is the wall time, in seconds, of the last timestep's
 * parallel steps.  log_csv appends a row (the caller formats it like the
 * FlameGPU backend, "%d"/"%f" joined by ',', src/backend/FlameGPUPrinter.cpp:
 * 226-245); rows are kept in memory and appended to the "hip.log_csv" file. */
int abl_hip_reduce_sum(abl_sim *sim, int agent_type, const char *member, double *out);
int abl_hip_reduce_count_eq(abl_sim *sim, int agent_type, const char *member, double value,
                            size_t *count);

/* sum(Type.member) exactly as abl_hip_reduce_sum computes it (same kernel,
 * same partial sums, same result bits), but allocating nothing: no
 * hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The partial sums live in a static 8 KiB device array, one per GPU (at
 * most 1024 blocks, each striding over the column, so any agent count
 * fits); the array's device address is looked up on the first call on each
 * device.  Calls on one GPU share the array and are serialised by a host
 * lock; a call waits for the sim's stream. */
int abl_hip_reduce_sum_noalloc(abl_sim *sim, int agent_type, const char *member, double *out);
int abl_hip_last_exec_time(abl_sim *sim, double *seconds);
int abl_hip_log_csv(abl_sim *sim, const char *row);
int abl_hip_log_text(abl_sim *sim, char *buf, size_t capacity, size_t *len);

/* ---- batched ensembles (no counterpart in the reference) -----------------
 * `members` independent runs of one program (uneven agent counts; parameters
 * per member, row-major members x num_params) advanced together: one counting
 * sort and one tiled step launch cover every member each timestep, and the
 * timestep is replayed as a captured HIP graph ("hip.graph", default on).
 * Member m's initialisation starts from abl_ens_member_seed(s0, s1, m); each
 * member's result is bit-identical to a run of it alone from that state with
 * the same "hip.cell_scale", "hip.tile", "hip.nlist", "hip.skin" and
 * "hip.list_cap".  Neighbour lists ("hip.nlist", default 1, when every member
 * has the same near radius): one set over all members' agents, and a rebuild
 * gate per member, so a member re-sorts and rebuilds its lists on exactly the
 * timesteps a run of it alone does.  Programs: boids2d. */
typedef struct abl_ensemble abl_ensemble;
const char *abl_ens_last_error(void);
int abl_ens_member_seed(uint64_t s0, uint64_t s1, int member, uint64_t out[2]);
int abl_ens_create(const char *program, int precision, int members, const int *counts,
                   int num_params, const char *const *param_names, const double *param_values,
                   abl_ensemble **out);
int abl_ens_set_config(abl_ensemble *ens, const char *key, const char *value);
int abl_ens_init(abl_ensemble *ens, uint64_t s0, uint64_t s1);
int abl_ens_simulate(abl_ensemble *ens, int timesteps);
/* waits; *ms receives the device time of the last simulate call */
int abl_ens_synchronize(abl_ensemble *ens, double *ms);
int abl_ens_member_count(abl_ensemble *ens, int member, size_t *count);
int abl_ens_agents(abl_ensemble *ens, size_t *total);
int abl_ens_download(abl_ensemble *ens, int member, void *aos, size_t capacity, size_t *count);
int abl_ens_destroy(abl_ensemble *ens);

/* ---- serving: continuous batching of runs (no counterpart in the
 * reference, which runs one simulation per process, src/main.cpp:280-311) --
 * Requests arrive at any time; a worker thread admits them in arrival order
 * while the estimated device footprint of the active runs fits `hbm_budget`
 * and at most `max_active` run at once, initialises each (serialised: the
 * generator is process-wide) from the request's generator state, and advances
 * every active run `chunk_steps` timesteps at a time on its own stream, so
 * runs overlap on the device.  Finished runs are downloaded (reference array
 * order) and their memory released.  Results do not depend on the mix. */
typedef struct abl_server abl_server;

typedef struct {
  const char *program;
  int precision;
  int num_params;
  const char *const *param_names;  /* -P overrides, derived constants folded */
  const double *param_values;
  int timesteps;
  uint64_t rng_state[2];           /* generator state the run's init starts from */
  size_t agents_hint;              /* most agents the run holds (admission) */
  size_t cells_hint;               /* grid cells of its environment (admission) */
  int num_config;                  /* -C overrides (abl_hip_set_config), or 0 */
  const char *const *config_keys;
  const char *const *config_values;
} abl_run_request;

enum { ABL_RUN_QUEUED = 0, ABL_RUN_ACTIVE = 1, ABL_RUN_DONE = 2, ABL_RUN_FAILED = 3 };

const char *abl_server_last_error(void);
int abl_server_create(size_t hbm_budget, int max_active, int chunk_steps, abl_server **out);
int abl_server_submit(abl_server *srv, const abl_run_request *req, uint64_t *run_id);
int abl_server_status(abl_server *srv, uint64_t run_id, int *state);
int abl_server_wait(abl_server *srv, uint64_t run_id);
int abl_server_result(abl_server *srv, uint64_t run_id, int agent_type, void *aos,
                      size_t capacity, size_t *count);
int abl_server_run_log(abl_server *srv, uint64_t run_id, char *buf, size_t capacity, size_t *len);
/* seconds: [0] queued (submit -> admit), [1] on device (admit -> done), [2] total */
int abl_server_run_times(abl_server *srv, uint64_t run_id, double out[3]);
/* [0] completed runs, [1] p50 and [2] p99 latency (submit -> result, s),
 * [3] agent-steps completed, [4] seconds from first admission to last
 * completion, [5] peak admitted footprint (bytes) */
int abl_server_stats(abl_server *srv, double out[6]);
int abl_server_destroy(abl_server *srv);

/* Device memory the sim holds, in bytes. */
int abl_hip_footprint(abl_sim *sim, size_t *bytes);

/* Timing of the last simulate call's kernels, in milliseconds, measured with
 * HIP events on the sim's stream: [0] whole call, [1] neighbour-step kernels,
 * [2] grid build (bin, scan, sort, gather).  Requires "hip.timing"="1". */
int abl_hip_last_timing(abl_sim *sim, double out_ms[3]);

#ifdef __cplusplus
}
#endif

#endif
