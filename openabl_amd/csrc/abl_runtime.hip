// abl_runtime.hip -- the -b hip runtime core: agent storage, the grid
// pipeline, the simulate loop and the C ABI declared in include/abl_hip.h.
//
// Storage.  Each agent type keeps one device array per scalar component of
// each member (struct-of-arrays; float2/float3 members split into x, y, z),
// two buffers per column plus a per-column "current" index, and an int32 id
// column (the agent's index in the reference's array).  A step function
// writes only the columns it assigns into their other buffer and flips those;
// columns it does not assign keep their values, which is the copy-through
// semantics of the Mason and FlameGPU backends and, for every model whose
// steps assign all members, of the C backend (src/backend/CPrinter.cpp:208-228).
//
// Grid pipeline (per near-loop agent type, rebuilt when its positions moved):
//   bin      key[i] = cell(pos[i]); count[key]++            (atomics, L2)
//   scan     start[c] = exclusive prefix of count           (wave64 scans)
//   scatter  slot = start[c] + --count[c]; tmp[slot] = i    (count returns to 0)
//   rank     stable placement inside each cell by id, fused with the gather
//            of every column into the other buffer          (HBM-bound)
// After the pipeline the agents are cell-sorted and, inside each cell, in
// ascending id, which is independent of history: a run gives the same result
// alone or beside other runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/abl_hip.h"
#include "abl_runtime.hpp"
#include "abl_pipeline.hpp"
#include "abl_gen_device.hpp"

using namespace abl;

// ---- errors -----------------------------------------------------------------

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(call)                                                            \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      return fail(ABL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                            \
  } while (0)

extern "C" const char *abl_hip_last_error(void) { return g_last_error.c_str(); }

// Device allocations of one call, freed when the call returns (error paths
// included).
struct ScopedDevice {
  std::vector<void *> ptrs;
  ScopedDevice() = default;
  ScopedDevice(const ScopedDevice &) = delete;
  ScopedDevice &operator=(const ScopedDevice &) = delete;
  ~ScopedDevice() {
    for (void *p : ptrs) (void) hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T **p, size_t bytes) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e == hipSuccess) ptrs.push_back(q);
    *p = (T *) q;
    return e;
  }
};

// ---- program registry -----------------------------------------------------

static const ProgramDesc *const *programs() {
  static const ProgramDesc *table[] = {
      program_circle(), program_circle3d(), program_boids2d(), program_game_of_life(),
      program_predator_prey(), nullptr,
  };
  return table;
}

// Programs registered at load time by generated step libraries
// (register_program), after the built-in ones; a registered program of an
// existing name replaces it for later lookups.
static std::mutex g_registry_mu;
static std::vector<const ProgramDesc *> &registered() {
  static std::vector<const ProgramDesc *> v;
  return v;
}

bool abl::register_program(const ProgramDesc *desc) {
  if (!desc || !desc->name) return false;
  std::lock_guard<std::mutex> g(g_registry_mu);
  registered().push_back(desc);
  return true;
}

const ProgramDesc *abl::find_program(const char *name) {
  if (!name) return nullptr;
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    for (auto it = registered().rbegin(); it != registered().rend(); ++it)
      if (std::strcmp((*it)->name, name) == 0) return *it;
  }
  for (int i = 0; programs()[i]; i++) {
    if (std::strcmp(programs()[i]->name, name) == 0) return programs()[i];
  }
  return nullptr;
}

extern "C" int abl_hip_program_count(void) {
  int n = 0;
  while (programs()[n]) n++;
  std::lock_guard<std::mutex> g(g_registry_mu);
  return n + (int) registered().size();
}

extern "C" const char *abl_hip_program_name(int index) {
  int n = 0;
  while (programs()[n]) n++;
  if (index >= 0 && index < n) return programs()[index]->name;
  std::lock_guard<std::mutex> g(g_registry_mu);
  if (index >= n && index < n + (int) registered().size()) return registered()[index - n]->name;
  return nullptr;
}

// ---- storage ---------------------------------------------------------------

namespace {

struct Column {
  int member;
  int comp;
  int elem;          // bytes per value on the device: 1 (bool), 4 (int, float), 8 (double)
  size_t aos_off;    // byte offset in the generated struct
  int aos_elem;      // bytes per value in the generated struct (8 with elem 4: a double
                     // held as float on the device, ABL_PREC_F32C)
};

struct AgentStore {
  const AgentDesc *desc = nullptr;
  std::vector<Column> cols;
  size_t aos_size = 0;
  int pos_col = -1;  // first position column
  int dim = 0;
  size_t n = 0, cap = 0;
  void *mem[2][kMaxCols] = {};
  int cur[kMaxCols] = {};
  int32_t *id[2] = {nullptr, nullptr};
  int id_cur = 0;
  // grid: per-agent keys and sort scratch, per-cell counts and (two) start
  // arrays, per-tile totals of the counts
  uint32_t *key = nullptr, *count = nullptr, *bsum = nullptr;
  uint32_t *startb[2] = {nullptr, nullptr};
  int start_cur = 0;
  uint32_t *tmp_idx = nullptr;
  int32_t *tmp_id = nullptr;
  size_t ncells = 0, cells_cap = 0;
  bool grid_valid = false;
  GridView gv{};
  double cell = 0.0;
  // delta re-sort (abl_pipeline.hpp): sorted keys of the stored order, mover
  // lists, and whether the stored order is the (cell, id) order of the last
  // build (same grid, same agents) with count/bsum/skey describing it
  uint32_t *skey[2] = {nullptr, nullptr};
  int skey_cur = 0;
  uint32_t *dirty = nullptr;
  unsigned long long *mhead = nullptr, *mnext = nullptr;
  uint32_t gen = 0;
  bool sorted = false;
  GridView sgv{};
  size_t sorted_n = 0;
  // neighbour lists (abl_tile.hpp): valid for the stored order while nl_ok
  bool nl_ok = false;
  uint16_t *nl_idx = nullptr, *nl_len = nullptr;
  uint32_t *nl_perm = nullptr;  // per own slot: the tile's agents in order of list rows
  uint4 *nl_prow0 = nullptr;    // per own slot: that agent's first list row
  double *nl_pos0 = nullptr;
  uint32_t *nl_gate = nullptr;  // [0..1] rebuild gates by timestep parity, [2] longest list,
                                // [4] gated rebuilds run,
                                // [8..15] work queue of the one-launch rebuild
  int nl_cap = 0;               // entries per agent
  size_t nl_ncap = 0;           // agents allocated
  // the rebuild rate of list mode (list_rate): a pinned copy of the device's
  // count of gated rebuilds, the event after that copy, the previous sample
  uint32_t *nl_host = nullptr;
  hipEvent_t nl_ev = nullptr;
  bool nl_pending = false;
  bool nl_off = false;          // lists given up for this type (rebuilt nearly every step)
  bool nl_kept = false;         // the rebuild rate was measured and the lists kept
  uint64_t nl_t_req = 0, nl_t_prev = 0;
  uint32_t nl_c_prev = 0;
  // run-time add/remove: ids stop being 0..n-1 once agents die or are born
  int64_t next_id = 0;
  bool ids_dense = true;
  void *bmem[kMaxCols] = {};  // newborn staging columns, indexed by parent slot
  size_t bcap = 0;
};

}  // namespace

struct abl_sim {
  const ProgramDesc *prog = nullptr;
  int precision = ABL_PREC_F64;
  std::vector<double> params;
  std::vector<AgentStore> types;
  hipStream_t stream = nullptr;
  int device = 0;
  int block = 256;
  int cell_scale = 1;
  int order = 0;
  int tile = 1;
  bool grid_delta = true;  // "hip.grid": delta re-sort when the stored order allows it
  float lds_slack = 2.0f;  // "hip.lds_slack": LDS tile records per expected halo agent
  bool nlist = true;       // "hip.nlist": neighbour lists for steps that support them
  bool nlist_set = false;  // "hip.nlist" given: lists kept whatever their rebuild rate
  double skin = kDefaultSkin;  // "hip.skin": list skin as a fraction of the near radius
  int list_cap = 0;        // "hip.list_cap": entries per list (0: sized from the data)
  double cell_inflate = 0.0;  // "hip.cell_inflate": cell widening factor (0: automatic)
  bool fused_rebuild = true;  // "hip.rebuild": gated re-sort + list build in one launch
  uint32_t barrier_spins = kPhaseSpins;   // "hip.barrier_spins": spin limit of its phase waits
  unsigned rebuild_blocks = 0;  // "hip.rebuild_blocks": its launch size (0: resident capacity)
  // "hip.defer_upload": abl_hip_upload keeps the records on the host (no
  // device call), and turning it off uploads what was kept
  bool defer_upload = false;
  std::vector<std::vector<char>> pending;  // per agent type
  std::vector<size_t> pending_n;
  std::vector<char> has_pending;
  bool timing = false;
  hipEvent_t ev[6] = {};
  double last_ms[3] = {0, 0, 0};
  // per-kernel timing ("hip.timing"): event pairs around every neighbour-step
  // launch (kind 1) and grid build (kind 2) of the current simulate call
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> spans;
  void *staging = nullptr;
  size_t staging_cap = 0;
  size_t footprint = 0;
  // add()/removeCurrent() scratch: per-agent flags, their scan, birth list
  uint8_t *dead = nullptr, *born = nullptr;
  uint32_t *fpos = nullptr, *fbsum = nullptr, *dscalar = nullptr;
  int32_t *bl_pid = nullptr;
  uint32_t *bl_slot = nullptr;
  size_t fcap = 0;
  uint64_t seed = 0x5EEDULL;
  uint64_t timestep = 0;
  // sequential step: log_csv rows and getLastExecTime
  std::string log;
  std::string log_path;
  double last_exec = 0.0;
};

// C struct layout of the generated agent struct (src/backend/CPrinter.cpp:250-253):
// members in order, each aligned to its scalar size, total padded to the
// largest alignment.
// ABL_PREC_F32C keeps the FP64 program's struct (host initialisation and
// save() unchanged) and holds, on the device, positions as double and every
// other float member as float.
static void layout_agent(AgentStore &s, const AgentDesc &d, int precision) {
  s.desc = &d;
  s.cols.clear();
  size_t off = 0, align_max = 1;
  const int fb = precision == ABL_PREC_F32 ? 4 : 8;
  for (size_t m = 0; m < d.members.size(); m++) {
    const MemberDesc &md = d.members[m];
    const int aelem = md.kind == K_BOOL ? 1 : md.kind == K_INT ? 4 : fb;
    const int elem = (precision == ABL_PREC_F32C && aelem == 8 && !md.is_pos) ? 4 : aelem;
    size_t a = (size_t) aelem;
    off = (off + a - 1) / a * a;
    int nc = kind_components(md.kind);
    if (md.is_pos) {
      s.pos_col = (int) s.cols.size();
      s.dim = nc;
    }
    for (int c = 0; c < nc; c++) {
      s.cols.push_back(Column{(int) m, c, elem, off + (size_t) c * aelem, aelem});
    }
    off += (size_t) nc * aelem;
    align_max = std::max(align_max, a);
  }
  s.aos_size = (off + align_max - 1) / align_max * align_max;
}

static void free_store(AgentStore &s) {
  for (int b = 0; b < 2; b++) {
    for (int c = 0; c < kMaxCols; c++) {
      if (s.mem[b][c]) (void) hipFree(s.mem[b][c]);
      s.mem[b][c] = nullptr;
    }
    if (s.id[b]) (void) hipFree(s.id[b]);
    s.id[b] = nullptr;
  }
  void *grid_ptrs[] = {s.key, s.count, s.startb[0], s.startb[1], s.bsum, s.tmp_idx, s.tmp_id,
                       s.skey[0], s.skey[1], s.dirty, s.mhead, s.mnext};
  for (void *p : grid_ptrs) {
    if (p) (void) hipFree(p);
  }
  s.key = s.count = s.startb[0] = s.startb[1] = s.bsum = s.tmp_idx = nullptr;
  s.tmp_id = nullptr;
  s.skey[0] = s.skey[1] = s.dirty = nullptr;
  s.mhead = s.mnext = nullptr;
  s.sorted = false;
  void *nl_ptrs[] = {s.nl_idx, s.nl_len, s.nl_pos0, s.nl_gate, s.nl_perm, s.nl_prow0};
  for (void *p : nl_ptrs) {
    if (p) (void) hipFree(p);
  }
  s.nl_idx = s.nl_len = nullptr;
  s.nl_perm = nullptr;
  s.nl_prow0 = nullptr;
  s.nl_pos0 = nullptr;
  s.nl_gate = nullptr;
  s.nl_cap = 0;
  s.nl_ncap = 0;
  s.nl_ok = false;
  if (s.nl_ev) (void) hipEventDestroy(s.nl_ev);
  if (s.nl_host) (void) hipHostFree(s.nl_host);
  s.nl_ev = nullptr;
  s.nl_host = nullptr;
  s.nl_pending = false;
  for (int c = 0; c < kMaxCols; c++) {
    if (s.bmem[c]) (void) hipFree(s.bmem[c]);
    s.bmem[c] = nullptr;
  }
  s.bcap = 0;
  s.n = s.cap = 0;
  s.ncells = s.cells_cap = 0;
  s.grid_valid = false;
}

// Bytes of one record in every column of a type (the device element sizes).
static size_t col_bytes(const AgentStore &s) {
  size_t b = 0;
  for (const Column &c : s.cols) b += (size_t) c.elem;
  return b;
}

// Neighbour-list bytes per agent besides the entries: length (2), lane order
// (4), first row (16), build positions (3 x 8).
constexpr size_t kListMetaBytes = 2 + 4 + 16 + 24;

// Device bytes one agent type holds: columns (two buffers), ids, grid and
// sort scratch, neighbour lists, newborn staging.
static size_t store_bytes(const AgentStore &s) {
  size_t b = 2 * s.cap * col_bytes(s);
  b += 2 * s.cap * 4;                      // ids
  b += s.cap * (4 + 4 + 4 + 8 + 8);        // key, tmp_idx, tmp_id, skey x2, mover next
  const size_t tiles = (s.cells_cap + 2047) / 2048;
  b += s.cells_cap * (4 + 8 + 4 + 8) + std::max<size_t>(tiles, 1) * 4;  // count, start x2, dirty, head; bsum
  if (s.nl_idx) b += s.nl_ncap * ((size_t) s.nl_cap * 2 + kListMetaBytes) + 16 * 4;
  b += s.bcap * col_bytes(s);              // newborn staging
  return b;
}

static int reserve_store(abl_sim *sim, AgentStore &s, size_t n) {
  if (n <= s.cap) return ABL_OK;
  size_t cap = std::max<size_t>(n, 1024);
  // a column grown to `cap` records, keeping its first n: the new buffer
  // replaces the old one only once the copy is done
  auto grow = [&](void **slot, size_t elem) -> int {
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, cap * elem));
    hipError_t e = hipSuccess;
    if (*slot && s.n) e = hipMemcpyAsync(p, *slot, s.n * elem, hipMemcpyDeviceToDevice, sim->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
    if (e != hipSuccess) {
      (void) hipFree(p);
      return fail(ABL_ERR_DEVICE, std::string("column copy: ") + hipGetErrorString(e));
    }
    if (*slot) HIP_TRY(hipFree(*slot));
    *slot = p;
    return ABL_OK;
  };
  for (int b = 0; b < 2; b++) {
    for (size_t c = 0; c < s.cols.size(); c++) {
      int rc = grow(&s.mem[b][c], (size_t) s.cols[c].elem);
      if (rc) return rc;
    }
    int rc = grow((void **) &s.id[b], 4);
    if (rc) return rc;
  }
  void *ptrs[] = {s.key, s.tmp_idx, s.tmp_id, s.skey[0], s.skey[1], s.mnext};
  for (void *p : ptrs) {
    if (p) HIP_TRY(hipFree(p));
  }
  HIP_TRY(hipMalloc(&s.key, cap * 4));
  HIP_TRY(hipMalloc(&s.tmp_idx, cap * 4));
  HIP_TRY(hipMalloc(&s.tmp_id, cap * 4));
  HIP_TRY(hipMalloc(&s.skey[0], cap * 4));
  HIP_TRY(hipMalloc(&s.skey[1], cap * 4));
  HIP_TRY(hipMalloc(&s.mnext, cap * 8));
  s.cap = cap;
  s.sorted = false;
  s.nl_ok = false;
  return ABL_OK;
}

static int reserve_staging(abl_sim *sim, size_t bytes) {
  if (bytes <= sim->staging_cap) return ABL_OK;
  if (sim->staging) HIP_TRY(hipFree(sim->staging));
  sim->staging = nullptr;
  HIP_TRY(hipMalloc(&sim->staging, bytes));
  sim->staging_cap = bytes;
  return ABL_OK;
}

// ---- kernels: transposition ------------------------------------------------

struct ColXfer {
  void *col[kMaxCols];
  uint32_t aos_off[kMaxCols];
  int elem[kMaxCols];
  int narrow[kMaxCols];  // 1: double in the struct, float on the device
  int ncols;
  uint32_t aos_size;
};

__global__ void k_aos_to_soa(const char *__restrict__ aos, ColXfer x, int32_t *__restrict__ id,
                             int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const char *rec = aos + (size_t) i * x.aos_size;
  for (int c = 0; c < x.ncols; c++) {
    const char *p = rec + x.aos_off[c];
    if (x.narrow[c]) {
      ((float *) x.col[c])[i] = (float) *(const double *) p;
      continue;
    }
    switch (x.elem[c]) {
      case 1: ((uint8_t *) x.col[c])[i] = *(const uint8_t *) p; break;
      case 4: ((uint32_t *) x.col[c])[i] = *(const uint32_t *) p; break;
      default: ((uint64_t *) x.col[c])[i] = *(const uint64_t *) p; break;
    }
  }
  id[i] = i;
}

// Write slot s back to record id[s] (un-permute to the reference's order),
// or to record s when id is null.
__global__ void k_soa_to_aos(char *__restrict__ aos, ColXfer x, const int32_t *__restrict__ id,
                             int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  char *rec = aos + (size_t) (id ? id[s] : s) * x.aos_size;
  for (int c = 0; c < x.ncols; c++) {
    char *p = rec + x.aos_off[c];
    if (x.narrow[c]) {
      *(double *) p = (double) ((const float *) x.col[c])[s];
      continue;
    }
    switch (x.elem[c]) {
      case 1: *(uint8_t *) p = ((const uint8_t *) x.col[c])[s]; break;
      case 4: *(uint32_t *) p = ((const uint32_t *) x.col[c])[s]; break;
      default: *(uint64_t *) p = ((const uint64_t *) x.col[c])[s]; break;
    }
  }
}

// ---- kernels: grid pipeline ----------------------------------------------------

template <typename P>
__global__ void k_bin(const P *__restrict__ px, const P *__restrict__ py, const P *__restrict__ pz,
                      GridView g, uint32_t *__restrict__ key, uint32_t *__restrict__ count, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int3 c = cell_of<P>(g, px[i], py[i], g.dim == 3 ? pz[i] : (P) 0);
  uint32_t k = ((uint32_t) c.z * (uint32_t) g.n[1] + (uint32_t) c.y) * (uint32_t) g.n[0] + (uint32_t) c.x;
  key[i] = k;
  atomicAdd(&count[k], 1u);
}

// ---- kernels: run-time removal and birth -----------------------------------------

__global__ void k_zero_u8(uint8_t *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

static __device__ __forceinline__ void copy_cols(const GatherCols &g, uint32_t src, uint32_t dst) {
  for (int c = 0; c < g.ncols; c++) {
    switch (g.elem[c]) {
      case 1: ((uint8_t *) g.dst[c])[dst] = ((const uint8_t *) g.src[c])[src]; break;
      case 4: ((uint32_t *) g.dst[c])[dst] = ((const uint32_t *) g.src[c])[src]; break;
      default: ((uint64_t *) g.dst[c])[dst] = ((const uint64_t *) g.src[c])[src]; break;
    }
  }
}

// Stable compaction after removeCurrent(): survivor s moves to pos[s], the
// exclusive prefix count of survivors before it.
__global__ void k_compact(const uint8_t *__restrict__ dead, const uint32_t *__restrict__ pos,
                          GatherCols g, const int32_t *__restrict__ id_in,
                          int32_t *__restrict__ id_out, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || dead[s]) return;
  const uint32_t d = pos[s];
  copy_cols(g, (uint32_t) s, d);
  id_out[d] = id_in[s];
}

// Births of one step: the k-th parent (in slot order) with born[s] set.
__global__ void k_birth_list(const uint8_t *__restrict__ born, const uint32_t *__restrict__ pos,
                             const int32_t *__restrict__ id, int32_t *__restrict__ pid,
                             uint32_t *__restrict__ slot, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !born[s]) return;
  pid[pos[s]] = id[s];
  slot[pos[s]] = (uint32_t) s;
}

// Append the newborns after the child type's agents in ascending parent id;
// they take ids next_id + rank.  A newborn's rank counts the smaller parent
// ids among all newborns, read 256 at a time through LDS (blockDim 256).
__global__ __launch_bounds__(256) void k_birth_append(const int32_t *__restrict__ pid,
                                                      const uint32_t *__restrict__ slot, int nb,
                                                      GatherCols g, int32_t *__restrict__ child_id,
                                                      int n_child, int32_t next_id) {
  __shared__ int32_t chunk[256];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t my = k < nb ? pid[k] : 0;
  uint32_t rank = 0;
  for (int base = 0; base < nb; base += 256) {  // every thread of the block takes part
    const int q = base + (int) threadIdx.x;
    if (q < nb) chunk[threadIdx.x] = pid[q];
    __syncthreads();
    const int m = nb - base < 256 ? nb - base : 256;
    for (int j = 0; j < m; j++) rank += chunk[j] < my ? 1u : 0u;
    __syncthreads();
  }
  if (k >= nb) return;
  const uint32_t dst = (uint32_t) n_child + rank;
  copy_cols(g, slot[k], dst);
  child_id[dst] = next_id + (int32_t) rank;
}

// ---- kernels: reductions for the sequential step -----------------------------------

// Per-block partial sums of a column (int64 for bool/int, double for float);
// the host adds the partials in block order, so the result is deterministic.
template <typename T, typename Acc>
__global__ __launch_bounds__(256) void k_reduce_sum(const T *__restrict__ col, int n,
                                                    Acc *__restrict__ partial) {
  __shared__ Acc lds[256];
  Acc acc = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc += (Acc) col[i];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int) threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

template <typename T>
__global__ __launch_bounds__(256) void k_reduce_count_eq(const T *__restrict__ col, int n, T v,
                                                         unsigned long long *__restrict__ partial) {
  __shared__ unsigned long long lds[256];
  unsigned long long acc = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc += col[i] == v ? 1 : 0;
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int) threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// ---- kernels: neighbour-set probe (parity) -----------------------------------

template <typename P>
__global__ void k_probe(const P *__restrict__ ax, const P *__restrict__ ay, const P *__restrict__ az,
                        const int32_t *__restrict__ aid, int n, const P *__restrict__ bx,
                        const P *__restrict__ by, const P *__restrict__ bz,
                        const int32_t *__restrict__ bid, GridView g, float t,
                        uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
                        int64_t *__restrict__ pairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P x = ax[i], y = ay[i], z = g.dim == 3 ? az[i] : (P) 0;
  const int3 c = cell_of<P>(g, x, y, z);
  uint32_t cnt = 0;
  uint32_t w = offsets ? offsets[i] : 0u;
  for_each_candidate(g, c.x, c.y, c.z, g.reach, [&](uint32_t j) {
    float f = g.dim == 3 ? dist_sq_f32<P>(bx[j], by[j], bz[j], x, y, z)
                         : dist_sq_f32<P>(bx[j], by[j], x, y);
    if (f > t) return;
    if (pairs) {
      pairs[2 * (size_t) w] = aid[i];
      pairs[2 * (size_t) w + 1] = bid[j];
      w++;
    }
    cnt++;
  });
  if (counts) counts[i] = cnt;
}

// ---- grid build ------------------------------------------------------------------

static void *col_cur(const AgentStore &s, int c) { return s.mem[s.cur[c]][c]; }
static void *col_other(const AgentStore &s, int c) { return s.mem[1 - s.cur[c]][c]; }

// Cells are widened by (1 + skin) for a type whose near loops walk neighbour
// lists, so the list radius (1 + skin) * r needs no more stencil reach than r
// does ("hip.cell_inflate" overrides the factor; 1 = none).
static double cell_inflation(const abl_sim *sim, const AgentStore &s) {
  if (sim->cell_inflate > 0.0) return sim->cell_inflate;
  if (!sim->nlist || sim->order != 0 || sim->tile == 0) return 1.0;
  for (const StepDesc &st : sim->prog->steps) {
    if (st.nlist && st.near_agent >= 0 && &sim->types[st.near_agent] == &s && !st.removes &&
        !st.adds)
      return 1.0 + sim->skin;
  }
  return 1.0;
}

static int setup_grid(abl_sim *sim, AgentStore &s) {
  double lo[3], hi[3], gran;
  sim->prog->env(sim->params.data(), lo, hi, &gran);
  if (!(gran > 0.0)) return fail(ABL_ERR_ARG, "environment granularity must be positive");
  const double cell =
      gran / (double) sim->cell_scale * (1.0 + 1.0 / 1048576.0) * cell_inflation(sim, s);
  size_t ncells = 1;
  GridView g{};
  g.dim = s.dim;
  for (int d = 0; d < 3; d++) {
    g.lo[d] = d < s.dim ? lo[d] : 0.0;
    int nd = 1;
    if (d < s.dim) {
      double span = hi[d] - lo[d];
      double q = std::ceil(span / cell);
      nd = q < 1.0 ? 1 : (int) q;
    }
    g.n[d] = nd;
    ncells *= (size_t) nd;
  }
  if (ncells > (size_t) 1 << 31) return fail(ABL_ERR_MEMORY, "grid has too many cells");
  g.inv_cell = 1.0 / cell;
  s.cell = cell;
  if (ncells + 1 > s.cells_cap) {
    void *ptrs[] = {s.count, s.startb[0], s.startb[1], s.bsum, s.dirty, s.mhead};
    for (void *p : ptrs) {
      if (p) HIP_TRY(hipFree(p));
    }
    s.count = s.startb[0] = s.startb[1] = s.bsum = s.dirty = nullptr;
    s.mhead = nullptr;
    HIP_TRY(hipMalloc(&s.count, (ncells + 1) * 4));
    HIP_TRY(hipMalloc(&s.startb[0], (ncells + 1) * 4));
    HIP_TRY(hipMalloc(&s.startb[1], (ncells + 1) * 4));
    size_t ntiles = (ncells + kScanTile - 1) / kScanTile;
    HIP_TRY(hipMalloc(&s.bsum, std::max<size_t>(ntiles, 1) * 4));
    HIP_TRY(hipMalloc(&s.dirty, (ncells + 1) * 4));
    HIP_TRY(hipMalloc(&s.mhead, (ncells + 1) * 8));
    HIP_TRY(hipMemsetAsync(s.count, 0, (ncells + 1) * 4, sim->stream));
    HIP_TRY(hipMemsetAsync(s.dirty, 0, (ncells + 1) * 4, sim->stream));
    HIP_TRY(hipMemsetAsync(s.mhead, 0, (ncells + 1) * 8, sim->stream));
    s.cells_cap = ncells + 1;
    s.sorted = false;
    s.nl_ok = false;
  }
  s.ncells = ncells;
  g.start = s.startb[s.start_cur];
  s.gv = g;
  return ABL_OK;
}

static bool same_grid(const GridView &a, const GridView &b) {
  for (int d = 0; d < 3; d++) {
    if (a.lo[d] != b.lo[d] || a.n[d] != b.n[d]) return false;
  }
  return a.inv_cell == b.inv_cell && a.dim == b.dim;
}

template <typename P>
static void launch_bin(abl_sim *sim, AgentStore &s, bool delta, int n,
                       const uint32_t *gate = nullptr) {
  const int pc = s.pos_col;
  const P *px = (const P *) col_cur(s, pc);
  const P *py = (const P *) col_cur(s, pc + 1);
  const P *pz = (const P *) (s.dim == 3 ? col_cur(s, pc + 2) : nullptr);
  if (delta) {
    MoverLists ml{s.dirty, s.mhead, s.mnext};
    hipLaunchKernelGGL(k_bin_delta<P>, grid_blocks(n, 256), 256, 0, sim->stream, px, py, pz, s.gv,
                       s.skey[s.skey_cur], s.key, s.count, s.bsum, ml, s.gen, n, gate);
  } else {
    hipLaunchKernelGGL(k_bin<P>, grid_blocks(n, 256), 256, 0, sim->stream, px, py, pz, s.gv, s.key,
                       s.count, n);
  }
}

// The delta re-sort's kernels into the other buffers (start_new, skey, ids,
// columns gc.dst); with a gate they do nothing unless *gate is set.
static void launch_delta(abl_sim *sim, AgentStore &s, const uint32_t *gate, uint32_t *start_old,
                         uint32_t *start_new, const GatherCols &gc) {
  const int n = (int) s.n;
  const int nc = (int) s.ncells;
  const int ntiles = (nc + kScanTile - 1) / kScanTile;
  hipStream_t st = sim->stream;
  if (n > 0) {
    if (s.cols[s.pos_col].elem == 8) launch_bin<double>(sim, s, true, n, gate);
    else launch_bin<float>(sim, s, true, n, gate);
  }
  hipLaunchKernelGGL(k_scan_apply, std::max(ntiles, 1), kScanBlock, 0, st, s.count, s.bsum,
                     start_new, nc, gate);
  if (n > 0) {
    MoverLists ml{s.dirty, s.mhead, s.mnext};
    hipLaunchKernelGGL(k_place, grid_blocks(n, 256), 256, 0, st, s.key, s.skey[s.skey_cur],
                       s.skey[1 - s.skey_cur], start_old, start_new, ml, s.gen, gc,
                       s.id[s.id_cur], s.id[1 - s.id_cur], n, gate);
  }
}

// Cell-sort the agents of `s` (stable by id inside a cell) and build the cell
// offsets.  Full build: bin with per-agent atomics, scan, scatter, rank by id
// and gather.  Delta build (the stored order is the (cell, id) order of the
// last build on the same grid): only agents that changed cell touch atomics,
// and one pass places every agent (abl_pipeline.hpp).  Both give the same
// order, a function of positions and ids only.
static int build_grid(abl_sim *sim, AgentStore &s) {
  if (s.pos_col < 0) return fail(ABL_ERR_ARG, "near loop over an agent type without position");
  int rc = setup_grid(sim, s);
  if (rc) return rc;
  const int n = (int) s.n;
  const int nc = (int) s.ncells;
  hipStream_t st = sim->stream;
  const bool delta = sim->grid_delta && n > 0 && s.mnext && s.sorted && s.sorted_n == s.n &&
                     same_grid(s.gv, s.sgv);
  s.gen += 1;
  if (s.gen == 0) s.gen = 1;  // tag 0 marks never-touched cells and lists
  const int ntiles = (nc + kScanTile - 1) / kScanTile;
  uint32_t *start_old = s.startb[s.start_cur];
  uint32_t *start_new = s.startb[1 - s.start_cur];
  const bool dbl = s.cols[s.pos_col].elem == 8;
  GatherCols gc{};
  gc.ncols = (int) s.cols.size();
  for (int c = 0; c < gc.ncols; c++) {
    gc.src[c] = col_cur(s, c);
    gc.dst[c] = col_other(s, c);
    gc.elem[c] = s.cols[c].elem;
  }
  if (delta) {
    launch_delta(sim, s, nullptr, start_old, start_new, gc);
  } else {
    // the counts of the previous build are still in `count`
    HIP_TRY(hipMemsetAsync(s.count, 0, (size_t) (nc + 1) * 4, st));
    if (n > 0) {
      if (dbl) launch_bin<double>(sim, s, false, n);
      else launch_bin<float>(sim, s, false, n);
    }
    hipLaunchKernelGGL(k_tile_totals, std::max(ntiles, 1), kScanBlock, 0, st, s.count, s.bsum, nc);
    hipLaunchKernelGGL(k_scan_apply, std::max(ntiles, 1), kScanBlock, 0, st, s.count, s.bsum,
                       start_new, nc);
    if (n > 0) {
      hipLaunchKernelGGL(k_scatter, grid_blocks(n, 256), 256, 0, st, s.key, s.count, start_new,
                         s.id[s.id_cur], s.tmp_idx, s.tmp_id, n);
      hipLaunchKernelGGL(k_rank_gather, grid_blocks(n, 256), 256, 0, st, s.tmp_idx, s.tmp_id,
                         s.key, start_new, gc, s.id[1 - s.id_cur], n, s.skey[1 - s.skey_cur],
                         s.count);
    }
  }
  HIP_TRY(hipGetLastError());
  if (n > 0) {
    for (int c = 0; c < gc.ncols; c++) s.cur[c] ^= 1;
    s.id_cur ^= 1;
    s.skey_cur ^= 1;
  }
  s.start_cur ^= 1;
  s.gv.start = start_new;
  s.sgv = s.gv;
  s.sorted = true;
  s.nl_ok = false;  // the stored order changed: lists index the old one
  s.sorted_n = s.n;
  s.grid_valid = true;
  return ABL_OK;
}

static hipEvent_t take_event(abl_sim *sim) {
  if (sim->evused == sim->evpool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    sim->evpool.push_back(e);
  }
  return sim->evpool[sim->evused++];
}

// Record the start of a timed span; returns the index to close it with.
static int span_open(abl_sim *sim, int kind) {
  if (!sim->timing) return -1;
  hipEvent_t a = take_event(sim), b = take_event(sim);
  if (!a || !b) return -1;
  (void) hipEventRecord(a, sim->stream);
  sim->spans.push_back({kind, {a, b}});
  return (int) sim->spans.size() - 1;
}

static void span_close(abl_sim *sim, int idx) {
  if (idx >= 0) (void) hipEventRecord(sim->spans[idx].second.second, sim->stream);
}

static int reach_for(const AgentStore &s, double radius) {
  double q = radius / s.cell * (1.0 + 1.0 / 2097152.0);
  int r = (int) std::ceil(q);
  return r < 1 ? 1 : r;
}

static double step_radius(const abl_sim *sim, const StepDesc &st) {
  if (st.radius_fn) return st.radius_fn(sim->precision);
  for (size_t p = 0; p < sim->prog->params.size(); p++) {
    if (std::strcmp(sim->prog->params[p].name, st.radius_param) == 0) {
      // under use_float the parameter is a float global, widened for the
      // double-literal product the near radius expression prints as
      double v = sim->params[p];
      if (sim->precision == ABL_PREC_F32) v = (double) (float) v;
      return st.radius_scale * v;
    }
  }
  return 0.0;
}

// ---- C ABI -----------------------------------------------------------------------

#define CHECK_SIM(sim) \
  if (!(sim)) return fail(ABL_ERR_ARG, "null simulation handle")


extern "C" int abl_hip_create(const char *program, int precision, abl_sim **out) {
  if (!program || !out) return fail(ABL_ERR_ARG, "null argument");
  *out = nullptr;
  if (precision != ABL_PREC_F64 && precision != ABL_PREC_F32 && precision != ABL_PREC_F32C)
    return fail(ABL_ERR_UNSUPPORTED, "unknown precision");
  const ProgramDesc *prog = find_program(program);
  if (!prog) return fail(ABL_ERR_ARG, std::string("unknown program ") + program);
  if (precision == ABL_PREC_F32C && !prog->f32c)
    return fail(ABL_ERR_UNSUPPORTED, std::string(program) + " has no FP32-compensated form");
  abl_sim *sim = new abl_sim();
  sim->prog = prog;
  sim->precision = precision;
  for (const ParamDesc &p : prog->params) sim->params.push_back(p.value);
  sim->types.resize(prog->agents.size());
  for (size_t t = 0; t < prog->agents.size(); t++) {
    layout_agent(sim->types[t], prog->agents[t], precision);
    if (sim->types[t].cols.size() > (size_t) kMaxCols) {
      delete sim;
      return fail(ABL_ERR_UNSUPPORTED, "agent type has too many member components");
    }
  }
  hipError_t e = hipGetDevice(&sim->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking);
  for (int k = 0; k < 6 && e == hipSuccess; k++) e = hipEventCreate(&sim->ev[k]);
  if (e != hipSuccess) {
    std::string msg = std::string("stream/event creation: ") + hipGetErrorString(e);
    delete sim;
    return fail(ABL_ERR_DEVICE, msg);
  }
  *out = sim;
  return ABL_OK;
}

extern "C" int abl_hip_destroy(abl_sim *sim) {
  CHECK_SIM(sim);
  (void) hipStreamSynchronize(sim->stream);
  for (AgentStore &s : sim->types) free_store(s);
  if (sim->staging) (void) hipFree(sim->staging);
  void *scratch[] = {sim->dead, sim->born, sim->fpos, sim->fbsum, sim->dscalar, sim->bl_pid,
                     sim->bl_slot};
  for (void *p : scratch) {
    if (p) (void) hipFree(p);
  }
  for (int k = 0; k < 6; k++) {
    if (sim->ev[k]) (void) hipEventDestroy(sim->ev[k]);
  }
  for (hipEvent_t e : sim->evpool) (void) hipEventDestroy(e);
  if (sim->stream) (void) hipStreamDestroy(sim->stream);
  delete sim;
  return ABL_OK;
}

extern "C" int abl_hip_set_param(abl_sim *sim, const char *name, double value) {
  CHECK_SIM(sim);
  if (!name) return fail(ABL_ERR_ARG, "null parameter name");
  for (size_t p = 0; p < sim->prog->params.size(); p++) {
    if (std::strcmp(sim->prog->params[p].name, name) == 0) {
      sim->params[p] = value;
      for (AgentStore &s : sim->types) {
        s.grid_valid = false;
        s.nl_ok = false;
      }
      return ABL_OK;
    }
  }
  return fail(ABL_ERR_ARG, std::string("Unknown parameter \"") + name + "\"");
}

extern "C" int abl_hip_get_param(abl_sim *sim, const char *name, double *value) {
  CHECK_SIM(sim);
  if (!name || !value) return fail(ABL_ERR_ARG, "null argument");
  for (size_t p = 0; p < sim->prog->params.size(); p++) {
    if (std::strcmp(sim->prog->params[p].name, name) == 0) {
      *value = sim->params[p];
      return ABL_OK;
    }
  }
  return fail(ABL_ERR_ARG, std::string("Unknown parameter \"") + name + "\"");
}

extern "C" int abl_hip_set_config(abl_sim *sim, const char *key, const char *value) {
  CHECK_SIM(sim);
  if (!key || !value) return fail(ABL_ERR_ARG, "null argument");
  std::string k(key), v(value);
  // launch settings decide the grid's cells and the lists' tiles, radius and
  // reach: rebuild both
  for (AgentStore &s : sim->types) {
    s.nl_ok = false;
    s.grid_valid = false;
  }
  if (k == "hip.block") {
    int b = std::atoi(value);
    if (b < 64 || b > 1024 || b % 64) return fail(ABL_ERR_ARG, "hip.block must be 64..1024, a multiple of 64");
    sim->block = b;
  } else if (k == "hip.cell_scale") {
    int c = std::atoi(value);
    if (c < 1 || c > 4) return fail(ABL_ERR_ARG, "hip.cell_scale must be 1..4");
    sim->cell_scale = c;
    for (AgentStore &s : sim->types) s.grid_valid = false;
  } else if (k == "hip.order") {
    if (v == "cell") sim->order = 0;
    else if (v == "index") sim->order = 1;
    else return fail(ABL_ERR_ARG, "hip.order must be cell or index");
  } else if (k == "hip.timing") {
    sim->timing = v == "1" || v == "true";
  } else if (k == "hip.tile") {
    int t = std::atoi(value);
    if (t < 0 || t > 3) return fail(ABL_ERR_ARG, "hip.tile must be 0..3");
    sim->tile = t;
  } else if (k == "hip.nlist") {
    if (v == "1" || v == "true") sim->nlist = true;
    else if (v == "0" || v == "false") sim->nlist = false;
    else return fail(ABL_ERR_ARG, "hip.nlist must be 0 or 1");
    sim->nlist_set = true;
  } else if (k == "hip.cell_inflate") {
    const double f = std::atof(value);
    if (!(f == 0.0 || (f >= 1.0 && f <= 2.0))) return fail(ABL_ERR_ARG, "hip.cell_inflate must be 0 or 1..2");
    sim->cell_inflate = f;
    for (AgentStore &s : sim->types) s.grid_valid = false;
  } else if (k == "hip.list_cap") {
    const int c = std::atoi(value);
    if (c < 0 || c > 65534) return fail(ABL_ERR_ARG, "hip.list_cap must be 0..65534");
    sim->list_cap = c;
  } else if (k == "hip.skin") {
    const double f = std::atof(value);
    if (!(f > 0.0 && f <= 1.0)) return fail(ABL_ERR_ARG, "hip.skin must be in (0, 1]");
    sim->skin = f;
  } else if (k == "hip.lds_slack") {
    const double f = std::atof(value);
    if (!(f >= 1.0 && f <= 4.0)) return fail(ABL_ERR_ARG, "hip.lds_slack must be 1..4");
    sim->lds_slack = (float) f;
  } else if (k == "hip.grid") {
    if (v == "delta") sim->grid_delta = true;
    else if (v == "full") sim->grid_delta = false;
    else return fail(ABL_ERR_ARG, "hip.grid must be delta or full");
  } else if (k == "hip.rebuild") {
    if (v == "fused") {
      // a fresh barrier state (an abandoned launch leaves its error flag set)
      for (AgentStore &s : sim->types) {
        if (s.nl_gate) {
          HIP_TRY(hipStreamSynchronize(sim->stream));
          HIP_TRY(hipMemset(s.nl_gate + 8, 0, 8 * 4));
        }
      }
      sim->fused_rebuild = true;
    } else if (v == "split") {
      sim->fused_rebuild = false;
    }
    else return fail(ABL_ERR_ARG, "hip.rebuild must be fused or split");
  } else if (k == "hip.barrier_spins") {
    const long long sp = std::atoll(value);
    if (sp < 0 || sp > (long long) kPhaseSpins) return fail(ABL_ERR_ARG, "hip.barrier_spins out of range");
    sim->barrier_spins = (uint32_t) sp;
  } else if (k == "hip.rebuild_blocks") {
    const long long nb = std::atoll(value);
    if (nb < 0 || nb > 0xFFF8) return fail(ABL_ERR_ARG, "hip.rebuild_blocks must be 0..65528");
    sim->rebuild_blocks = (unsigned) nb;
  } else if (k == "hip.defer_upload") {
    if (v == "1") {
      sim->defer_upload = true;
      sim->pending.assign(sim->types.size(), {});
      sim->pending_n.assign(sim->types.size(), 0);
      sim->has_pending.assign(sim->types.size(), 0);
    } else if (v == "0") {
      sim->defer_upload = false;
      for (size_t t = 0; t < sim->has_pending.size(); t++) {
        if (!sim->has_pending[t]) continue;
        const int rc = abl_hip_upload(sim, (int) t, sim->pending[t].data(), sim->pending_n[t]);
        if (rc) return rc;
      }
      sim->pending.clear();
      sim->pending_n.clear();
      sim->has_pending.clear();
    } else {
      return fail(ABL_ERR_ARG, "hip.defer_upload must be 0 or 1");
    }
  } else if (k == "hip.seed") {
    sim->seed = std::strtoull(value, nullptr, 0);
  } else if (k == "hip.log_csv") {
    sim->log_path = v;
  } else {
    return fail(ABL_ERR_ARG, "unknown config key " + k);
  }
  return ABL_OK;
}

extern "C" int abl_hip_agent_type_count(abl_sim *sim) {
  if (!sim) return fail(-1, "null simulation handle");
  return (int) sim->types.size();
}

extern "C" const char *abl_hip_agent_type_name(abl_sim *sim, int t) {
  if (!sim || t < 0 || t >= (int) sim->types.size()) return nullptr;
  return sim->types[t].desc->name;
}

extern "C" size_t abl_hip_agent_size(abl_sim *sim, int t) {
  if (!sim || t < 0 || t >= (int) sim->types.size()) return 0;
  return sim->types[t].aos_size;
}

static ColXfer xfer_of(const AgentStore &s, bool current) {
  ColXfer x{};
  x.ncols = (int) s.cols.size();
  x.aos_size = (uint32_t) s.aos_size;
  for (int c = 0; c < x.ncols; c++) {
    x.col[c] = current ? col_cur(s, c) : col_other(s, c);
    x.aos_off[c] = (uint32_t) s.cols[c].aos_off;
    x.elem[c] = s.cols[c].elem;
    x.narrow[c] = s.cols[c].aos_elem != s.cols[c].elem ? 1 : 0;
  }
  return x;
}

extern "C" int abl_hip_upload(abl_sim *sim, int t, const void *aos, size_t count) {
  CHECK_SIM(sim);
  if (t < 0 || t >= (int) sim->types.size()) return fail(ABL_ERR_ARG, "agent type out of range");
  if (count && !aos) return fail(ABL_ERR_ARG, "null agent data");
  if (count > (size_t) INT32_MAX) return fail(ABL_ERR_ARG, "too many agents");
  AgentStore &s = sim->types[t];
  if (sim->defer_upload) {
    // kept on the host until hip.defer_upload is turned off (no device call)
    const char *src = static_cast<const char *>(aos);
    sim->pending[(size_t) t].assign(src, src + count * s.aos_size);
    sim->pending_n[(size_t) t] = count;
    sim->has_pending[(size_t) t] = 1;
    return ABL_OK;
  }
  int rc = reserve_store(sim, s, count);
  if (rc) return rc;
  s.n = count;
  s.grid_valid = false;
  s.sorted = false;
  s.nl_ok = false;
  s.ids_dense = true;
  s.next_id = (int64_t) count;
  if (count == 0) return ABL_OK;
  rc = reserve_staging(sim, count * s.aos_size);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(sim->staging, aos, count * s.aos_size, hipMemcpyHostToDevice, sim->stream));
  ColXfer x = xfer_of(s, true);
  hipLaunchKernelGGL(k_aos_to_soa, grid_blocks((int) count, 256), 256, 0, sim->stream,
                     (const char *) sim->staging, x, s.id[s.id_cur], (int) count);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(sim->stream));
  return ABL_OK;
}

extern "C" int abl_hip_count(abl_sim *sim, int t, size_t *count) {
  CHECK_SIM(sim);
  if (t < 0 || t >= (int) sim->types.size() || !count) return fail(ABL_ERR_ARG, "bad argument");
  *count = sim->types[t].n;
  return ABL_OK;
}

static int check_rebuild(abl_sim *sim);

extern "C" int abl_hip_download(abl_sim *sim, int t, void *aos, size_t capacity, size_t *count) {
  CHECK_SIM(sim);
  if (t < 0 || t >= (int) sim->types.size()) return fail(ABL_ERR_ARG, "agent type out of range");
  AgentStore &s = sim->types[t];
  if (count) *count = s.n;
  if (s.n == 0) return ABL_OK;
  if (!aos || capacity < s.n) return fail(ABL_ERR_ARG, "download buffer too small");
  HIP_TRY(hipStreamSynchronize(sim->stream));
  int rc = check_rebuild(sim);
  if (rc) return rc;
  rc = reserve_staging(sim, s.n * s.aos_size);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(sim->staging, 0, s.n * s.aos_size, sim->stream));
  ColXfer x = xfer_of(s, true);
  if (s.ids_dense) {
    // ids are 0..n-1: un-permute on the device
    hipLaunchKernelGGL(k_soa_to_aos, grid_blocks((int) s.n, 256), 256, 0, sim->stream,
                       (char *) sim->staging, x, s.id[s.id_cur], (int) s.n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(aos, sim->staging, s.n * s.aos_size, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    return ABL_OK;
  }
  // after add/remove: records in slot order, then ordered by id on the host
  hipLaunchKernelGGL(k_soa_to_aos, grid_blocks((int) s.n, 256), 256, 0, sim->stream,
                     (char *) sim->staging, x, (const int32_t *) nullptr, (int) s.n);
  HIP_TRY(hipGetLastError());
  std::vector<char> recs(s.n * s.aos_size);
  std::vector<int32_t> ids(s.n);
  HIP_TRY(hipMemcpyAsync(recs.data(), sim->staging, recs.size(), hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(hipMemcpyAsync(ids.data(), s.id[s.id_cur], s.n * 4, hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(hipStreamSynchronize(sim->stream));
  std::vector<uint32_t> order(s.n);
  for (size_t i = 0; i < s.n; i++) order[i] = (uint32_t) i;
  std::sort(order.begin(), order.end(), [&](uint32_t p, uint32_t q) { return ids[p] < ids[q]; });
  for (size_t i = 0; i < s.n; i++) {
    std::memcpy((char *) aos + i * s.aos_size, recs.data() + (size_t) order[i] * s.aos_size, s.aos_size);
  }
  return ABL_OK;
}

// Ids of the agents of one type in download order (ascending).
extern "C" int abl_hip_download_ids(abl_sim *sim, int t, int64_t *ids, size_t capacity,
                                    size_t *count) {
  CHECK_SIM(sim);
  if (t < 0 || t >= (int) sim->types.size()) return fail(ABL_ERR_ARG, "agent type out of range");
  AgentStore &s = sim->types[t];
  if (count) *count = s.n;
  if (s.n == 0) return ABL_OK;
  if (!ids || capacity < s.n) return fail(ABL_ERR_ARG, "id buffer too small");
  std::vector<int32_t> h(s.n);
  HIP_TRY(hipMemcpyAsync(h.data(), s.id[s.id_cur], s.n * 4, hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(hipStreamSynchronize(sim->stream));
  std::sort(h.begin(), h.end());
  for (size_t i = 0; i < s.n; i++) ids[i] = h[i];
  return ABL_OK;
}

static int agent_index(abl_sim *sim, const char *name) {
  for (size_t t = 0; t < sim->types.size(); t++) {
    if (std::strcmp(sim->types[t].desc->name, name) == 0) return (int) t;
  }
  return -1;
}

extern "C" int abl_hip_upload_all(abl_sim *sim, const void *agents, const abl_agent_info *info) {
  CHECK_SIM(sim);
  if (!agents || !info) return fail(ABL_ERR_ARG, "null argument");
  for (const abl_agent_info *a = info; a->name; a++) {
    int t = agent_index(sim, a->name);
    if (t < 0) return fail(ABL_ERR_ARG, std::string("unknown agent type ") + a->name);
    const abl_type_info *end = a->info;
    while (end->type != ABL_TYPE_END) end++;
    if (end->offset != sim->types[t].aos_size)
      return fail(ABL_ERR_ARG, std::string("struct size mismatch for ") + a->name);
    const abl_dyn_array *arr = (const abl_dyn_array *) ((const char *) agents + a->offset);
    int rc = abl_hip_upload(sim, t, arr->values, arr->len);
    if (rc) return rc;
  }
  return ABL_OK;
}

extern "C" int abl_hip_download_all(abl_sim *sim, void *agents, const abl_agent_info *info) {
  CHECK_SIM(sim);
  if (!agents || !info) return fail(ABL_ERR_ARG, "null argument");
  for (const abl_agent_info *a = info; a->name; a++) {
    int t = agent_index(sim, a->name);
    if (t < 0) return fail(ABL_ERR_ARG, std::string("unknown agent type ") + a->name);
    abl_dyn_array *arr = (abl_dyn_array *) ((char *) agents + a->offset);
    size_t n = sim->types[t].n;
    if (arr->cap < n) {
      void *p = std::realloc(arr->values, n * sim->types[t].aos_size);
      if (!p) return fail(ABL_ERR_MEMORY, "host allocation failed");
      arr->values = p;
      arr->cap = n;
    }
    int rc = abl_hip_download(sim, t, arr->values, arr->cap, &arr->len);
    if (rc) return rc;
  }
  return ABL_OK;
}

extern "C" int abl_hip_init_agents(abl_sim *sim) {
  CHECK_SIM(sim);
  if (!sim->prog->host_init) return fail(ABL_ERR_UNSUPPORTED, "program has no host initialisation");
  return sim->prog->host_init(sim, sim->params.data(), sim->precision);
}

// ---- run-time add()/removeCurrent() ------------------------------------------------

static int reserve_flags(abl_sim *sim, size_t n) {
  if (n <= sim->fcap && sim->dead) return ABL_OK;
  size_t cap = std::max<size_t>(n, 1024);
  void *ptrs[] = {sim->dead, sim->born, sim->fpos, sim->fbsum, sim->bl_pid, sim->bl_slot};
  HIP_TRY(hipStreamSynchronize(sim->stream));
  for (void *p : ptrs) {
    if (p) HIP_TRY(hipFree(p));
  }
  HIP_TRY(hipMalloc(&sim->dead, cap));
  HIP_TRY(hipMalloc(&sim->born, cap));
  HIP_TRY(hipMalloc(&sim->fpos, (cap + 1) * 4));
  HIP_TRY(hipMalloc(&sim->fbsum, ((cap + kScanTile - 1) / kScanTile + 1) * 4));
  HIP_TRY(hipMalloc(&sim->bl_pid, cap * 4));
  HIP_TRY(hipMalloc(&sim->bl_slot, cap * 4));
  sim->fcap = cap;
  return ABL_OK;
}

// Exclusive scan of n flags into sim->fpos (INV: of 1 - flag); returns the
// total through *total (one read-back).
template <bool INV>
static int scan_flags(abl_sim *sim, const uint8_t *flags, int n, uint32_t *total) {
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  hipStream_t st = sim->stream;
  hipLaunchKernelGGL((k_scan_tiles<uint8_t, INV>), ntiles, kScanBlock, 0, st, flags, sim->fpos,
                     sim->fbsum, n);
  hipLaunchKernelGGL(k_scan_totals, 1, 1024, 0, st, sim->fbsum, ntiles, sim->fpos, n);
  hipLaunchKernelGGL(k_scan_add, grid_blocks(n, 256), 256, 0, st, sim->fpos, sim->fbsum, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(total, sim->fpos + n, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return ABL_OK;
}

static int reserve_births(abl_sim *sim, AgentStore &child, size_t parents) {
  if (parents <= child.bcap && child.bmem[0]) return ABL_OK;
  size_t cap = std::max<size_t>(parents, 1024);
  HIP_TRY(hipStreamSynchronize(sim->stream));
  for (size_t c = 0; c < child.cols.size(); c++) {
    if (child.bmem[c]) HIP_TRY(hipFree(child.bmem[c]));
    child.bmem[c] = nullptr;
    HIP_TRY(hipMalloc(&child.bmem[c], cap * (size_t) child.cols[c].elem));
  }
  child.bcap = cap;
  return ABL_OK;
}

// Drop the agents whose removeCurrent() flag is set, keeping order.
static int apply_removals(abl_sim *sim, AgentStore &a, int n) {
  uint32_t alive = 0;
  int rc = scan_flags<true>(sim, sim->dead, n, &alive);
  if (rc) return rc;
  if ((int) alive == n) return ABL_OK;
  GatherCols g{};
  g.ncols = (int) a.cols.size();
  for (int c = 0; c < g.ncols; c++) {
    g.src[c] = col_cur(a, c);
    g.dst[c] = col_other(a, c);
    g.elem[c] = a.cols[c].elem;
  }
  hipLaunchKernelGGL(k_compact, grid_blocks(n, 256), 256, 0, sim->stream, sim->dead, sim->fpos, g,
                     a.id[a.id_cur], a.id[1 - a.id_cur], n);
  HIP_TRY(hipGetLastError());
  for (int c = 0; c < g.ncols; c++) a.cur[c] ^= 1;
  a.id_cur ^= 1;
  a.n = alive;
  a.ids_dense = false;
  a.grid_valid = false;
  a.sorted = false;
  a.nl_ok = false;
  return ABL_OK;
}

// Append the step's newborns (staged per parent slot) to the child type.
static int apply_births(abl_sim *sim, AgentStore &parent, AgentStore &child, int n_parent) {
  uint32_t nb = 0;
  int rc = scan_flags<false>(sim, sim->born, n_parent, &nb);
  if (rc) return rc;
  if (nb == 0) return ABL_OK;
  hipLaunchKernelGGL(k_birth_list, grid_blocks(n_parent, 256), 256, 0, sim->stream, sim->born,
                     sim->fpos, parent.id[parent.id_cur], sim->bl_pid, sim->bl_slot, n_parent);
  GatherCols g{};
  g.ncols = (int) child.cols.size();
  for (int c = 0; c < g.ncols; c++) {
    g.src[c] = child.bmem[c];
    g.dst[c] = col_cur(child, c);
    g.elem[c] = child.cols[c].elem;
  }
  hipLaunchKernelGGL(k_birth_append, grid_blocks((int) nb, 256), 256, 0, sim->stream, sim->bl_pid,
                     sim->bl_slot, (int) nb, g, child.id[child.id_cur], (int) child.n,
                     (int32_t) child.next_id);
  HIP_TRY(hipGetLastError());
  child.n += nb;
  child.next_id += nb;
  child.ids_dense = false;
  child.grid_valid = false;
  child.sorted = false;
  child.nl_ok = false;
  return ABL_OK;
}

// ---- neighbour lists ---------------------------------------------------------------
// A step whose near loop runs over its own agent type, with a launcher that
// supports them, walks per-agent neighbour lists (abl_tile.hpp) instead of
// the stencil.  The lists are rebuilt on the device -- delta re-sort into the
// other buffers, list build, copy back -- on timesteps whose gate the
// previous step set, so the host never waits and its buffer bookkeeping does
// not depend on whether a rebuild ran.

static bool list_eligible(const abl_sim *sim, const StepDesc &st, const AgentStore &a) {
  return sim->nlist && st.nlist && st.near_agent == st.agent && sim->order == 0 && sim->tile != 0 &&
         !st.removes && !st.adds && a.n > 0 && a.pos_col >= 0 && a.n < (size_t) INT32_MAX &&
         !a.nl_off;
}

// With "hip.nlist" left at its default, lists are given up for an agent type
// whose gated rebuild runs on most timesteps (its agents outrun the skin
// every step): a re-sort and list build per timestep then costs more than
// the stencil walk, which gives the same results.  The device counts the
// rebuilds it runs (NListArgs::stats[2]; the list entry reads the count
// while it waits anyway).  kListSample timesteps later the host copies the
// count to pinned memory behind an event, and it reads the copy when it has
// issued another kListSample timesteps (waiting for the event if need be:
// the device still has those timesteps queued, so the wait costs it
// nothing).  The decision is thus taken at a fixed timestep, whatever the
// host's and the device's pace, and runs give the same bits alone, batched
// or served.  It is taken once per type: a type that stops outrunning its
// skin later keeps the stencil walk.
constexpr uint64_t kListSample = 16;
static int list_rate(abl_sim *sim, AgentStore &b) {
  if (sim->nlist_set || b.nl_off || b.nl_kept) return ABL_OK;
  if (!b.nl_host) {
    HIP_TRY(hipHostMalloc((void **) &b.nl_host, 4, hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&b.nl_ev, hipEventDisableTiming));
  }
  if (b.nl_pending) {
    if (sim->timestep - b.nl_t_req < kListSample) return ABL_OK;
    HIP_TRY(hipEventSynchronize(b.nl_ev));
    const uint32_t c = *b.nl_host;
    const uint64_t dt = b.nl_t_req - b.nl_t_prev;
    if (dt >= kListSample && 2 * (uint64_t) (c - b.nl_c_prev) > dt) {
      b.nl_off = true;
      b.sorted = false;  // the stencil walk starts from a full sort
    } else {
      b.nl_kept = true;
    }
    b.nl_c_prev = c;
    b.nl_t_prev = b.nl_t_req;
    b.nl_pending = false;
    return ABL_OK;
  }
  if (sim->timestep - b.nl_t_prev >= kListSample) {
    HIP_TRY(hipMemcpyAsync(b.nl_host, b.nl_gate + 4, 4, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipEventRecord(b.nl_ev, sim->stream));
    b.nl_t_req = sim->timestep;
    b.nl_pending = true;
  }
  return ABL_OK;
}

static int reserve_lists(abl_sim *sim, AgentStore &s, int cap) {
  cap = (cap + 7) & ~7;  // whole 16-byte rows of entries
  if (cap > 65528) cap = 65528;
  if (s.nl_idx && s.nl_cap >= cap && s.nl_ncap >= s.n) return ABL_OK;
  void *ptrs[] = {s.nl_idx, s.nl_len, s.nl_pos0, s.nl_perm, s.nl_prow0};
  HIP_TRY(hipStreamSynchronize(sim->stream));
  for (void *p : ptrs) {
    if (p) HIP_TRY(hipFree(p));
  }
  const size_t n = std::max(s.n, s.cap);
  HIP_TRY(hipMalloc(&s.nl_idx, (size_t) cap * n * 2));
  HIP_TRY(hipMalloc(&s.nl_len, n * 2));
  HIP_TRY(hipMalloc(&s.nl_pos0, 3 * n * 8));
  HIP_TRY(hipMalloc(&s.nl_perm, n * 4));
  HIP_TRY(hipMalloc(&s.nl_prow0, n * 16));
  if (!s.nl_gate) {
    HIP_TRY(hipMalloc(&s.nl_gate, 16 * 4));
    HIP_TRY(hipMemset(s.nl_gate, 0, 16 * 4));
  }
  s.nl_cap = cap;
  s.nl_ncap = n;
  return ABL_OK;
}

static NListArgs list_args(const abl_sim *sim, const AgentStore &s, double r, double rl) {
  NListArgs nl{};
  nl.idx = s.nl_idx;
  nl.len = s.nl_len;
  nl.perm = s.nl_perm;
  nl.prow0 = s.nl_prow0;
  nl.pos0 = s.nl_pos0;
  nl.gate = s.nl_gate;
  nl.stats = s.nl_gate + 2;
  nl.n = (int) s.n;
  nl.cap = s.nl_cap;
  nl.stride = (int) s.nl_ncap;
  nl.parity = (int) (sim->timestep & 1u);
  nl.rc_list = (float) (rl / s.cell * (1.0 + 1.0 / 1048576.0));
  // widened far past the float error of the records (relative records: ~1e-6)
  const double rw = rl * (1.0 + 1e-6) + 1e-5 * r;
  nl.r2_list = rw * rw;
  const double h = 0.5 * (rl - r);
  nl.half_skin2 = h * h;
  nl.ncells = (int) s.ncells;
  return nl;
}

// Enter list mode on a fresh grid: build the lists from the current buffers,
// grow them once if the longest list did not fit (one host wait per entry).
static int enter_lists(abl_sim *sim, AgentStore &b, const StepDesc &st, const StepLaunch &Lb,
                       double r, double rl) {
  const double vol = b.dim == 3 ? 4.0 / 3.0 * M_PI * rl * rl * rl : M_PI * rl * rl;
  const double env = (double) b.ncells * std::pow(b.cell, (double) b.dim);
  const double expect = env > 0 ? vol * (double) b.n / env : 0.0;
  const int want = sim->list_cap ? sim->list_cap : std::max(b.nl_cap, (int) (1.5 * expect) + 16);
  if (sim->list_cap && b.nl_cap != ((sim->list_cap + 7) & ~7)) b.nl_cap = 0;  // size requested
  int rc = reserve_lists(sim, b, want);
  if (rc) return rc;
  for (int attempt = 0; attempt < 2; attempt++) {
    HIP_TRY(hipMemsetAsync(b.nl_gate, 0, 4 * 4, sim->stream));
    NListArgs nl = list_args(sim, b, r, rl);
    StepLaunch L = Lb;
    L.nl_mode = 1;
    L.nl = &nl;
    st.launch(L);
    HIP_TRY(hipGetLastError());
    uint32_t words[3] = {0, 0, 0};  // longest list; -; gated rebuilds run so far
    HIP_TRY(hipMemcpyAsync(words, b.nl_gate + 2, sizeof(words), hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    const uint32_t longest = words[0];
    // the rebuild rate is measured from here (list_rate)
    b.nl_c_prev = words[2];
    b.nl_t_prev = sim->timestep;
    b.nl_pending = false;
    // grow to the longest list, but to at most four times the estimate: an
    // agent in a dense cluster with a longer list walks the stencil instead
    const int grow = std::min((int) (1.25 * longest) + 16, 4 * want);
    if ((int) longest <= b.nl_cap || grow <= b.nl_cap || b.nl_cap >= 65528 || sim->list_cap) break;
    rc = reserve_lists(sim, b, grow);
    if (rc) return rc;
  }
  b.nl_ok = true;
  return ABL_OK;
}

// The gated rebuild of one timestep: re-sort into the other buffers, build
// the lists there and copy the agents back, or nothing when the gate is clear.
static int gated_rebuild(abl_sim *sim, AgentStore &b, const StepDesc &st, const StepLaunch &Lb,
                         double r, double rl) {
  b.gen += 1;
  if (b.gen == 0) b.gen = 1;
  const uint32_t *gate = b.nl_gate + (sim->timestep & 1u);
  uint32_t *start_cur = b.startb[b.start_cur];
  uint32_t *start_other = b.startb[1 - b.start_cur];
  GatherCols placed{};
  placed.ncols = (int) b.cols.size();
  for (int c = 0; c < placed.ncols; c++) {
    placed.src[c] = col_cur(b, c);
    placed.dst[c] = col_other(b, c);
    placed.elem[c] = b.cols[c].elem;
  }
  // the re-sort in one launch with the list build (tiled_rebuild), when the
  // program's launch function can run it; else four gated launches
  DeltaArgs d{};
  int launched = 0;
  if (sim->fused_rebuild && b.n > 0 && b.mnext) {
    const int pc = b.pos_col;
    d.px = col_cur(b, pc);
    d.py = col_cur(b, pc + 1);
    d.pz = b.dim == 3 ? col_cur(b, pc + 2) : nullptr;
    d.pos_elem = b.cols[pc].elem;
    d.g = b.gv;
    d.skey = b.skey[b.skey_cur];
    d.key = b.key;
    d.count = b.count;
    d.bsum = b.bsum;
    d.ml = MoverLists{b.dirty, b.mhead, b.mnext};
    d.gen = b.gen;
    d.n = (int) b.n;
    d.start_old = start_cur;
    d.start_new = start_other;
    d.ncells = (int) b.ncells;
    d.nscan = std::max(((int) b.ncells + kScanTile - 1) / kScanTile, 1);
    d.skey_out = b.skey[1 - b.skey_cur];
    d.gc = placed;
    d.id_in = b.id[b.id_cur];
    d.id_out = b.id[1 - b.id_cur];
    d.q = reinterpret_cast<RebuildQueue *>(b.nl_gate + 8);
    d.spins = sim->barrier_spins;
    d.blocks = sim->rebuild_blocks;
    d.launched = &launched;
  } else {
    launch_delta(sim, b, gate, start_cur, start_other, placed);
  }
  NListArgs nl = list_args(sim, b, r, rl);
  nl.gated = 1;
  nl.copy.ncols = placed.ncols;
  for (int c = 0; c < placed.ncols; c++) {
    nl.copy.src[c] = placed.dst[c];
    nl.copy.dst[c] = (void *) placed.src[c];
    nl.copy.elem[c] = placed.elem[c];
  }
  nl.id_src = b.id[1 - b.id_cur];
  nl.id_dst = b.id[b.id_cur];
  nl.skey_src = b.skey[1 - b.skey_cur];
  nl.skey_dst = b.skey[b.skey_cur];
  nl.start_src = start_other;
  nl.start_dst = start_cur;
  StepLaunch L = Lb;
  for (size_t c = 0; c < b.cols.size(); c++) L.in.c[c] = col_other(b, (int) c);
  L.grid.start = start_other;
  L.nl = &nl;
  if (d.launched) {
    L.nl_mode = 3;
    L.delta = &d;
    st.launch(L);
    HIP_TRY(hipGetLastError());
    if (launched) return ABL_OK;
    launch_delta(sim, b, gate, start_cur, start_other, placed);
  }
  L.nl_mode = 1;
  L.delta = nullptr;
  st.launch(L);
  HIP_TRY(hipGetLastError());
  return ABL_OK;
}

// A one-launch rebuild whose phase wait ran past its spin limit left its
// agent type half re-sorted.  Every wait of the work queue ends by
// construction (abl_pipeline.hpp), so this happens only with a spin limit
// forced far below a phase's duration ("hip.barrier_spins"); it is reported at
// the next synchronisation as a device error, and the runtime rebuilds from
// scratch with the split launches from then on.
static int check_rebuild(abl_sim *sim) {
  for (AgentStore &s : sim->types) {
    if (!s.nl_gate) continue;
    uint32_t err = 0;
    HIP_TRY(hipMemcpy(&err, s.nl_gate + 8 + offsetof(RebuildQueue, error) / 4, 4,
                      hipMemcpyDeviceToHost));
    if (err) {
      HIP_TRY(hipMemset(s.nl_gate + 8, 0, 8 * 4));
      sim->fused_rebuild = false;
      for (AgentStore &t : sim->types) {
        t.nl_ok = false;
        t.grid_valid = false;
        t.sorted = false;
      }
      return fail(ABL_ERR_DEVICE, "one-launch neighbour-list rebuild ran past its spin limit; "
                                  "results since the last synchronisation are invalid");
    }
  }
  return ABL_OK;
}

static int launch_step(abl_sim *sim, const StepDesc &st, uint32_t ordinal) {
  AgentStore &a = sim->types[st.agent];
  if (st.removes || st.adds) {
    int rc = reserve_flags(sim, 2 * a.n);
    if (rc) return rc;
    if (st.adds) {
      AgentStore &c = sim->types[st.adds - 1];
      rc = reserve_store(sim, c, c.n + a.n);
      if (rc) return rc;
      rc = reserve_births(sim, c, a.n);
      if (rc) return rc;
    }
  }
  StepLaunch L{};
  L.stream = sim->stream;
  L.precision = sim->precision;
  L.n = (int) a.n;
  L.params = sim->params.data();
  L.block = sim->block;
  L.order = sim->order;
  for (size_t c = 0; c < a.cols.size(); c++) {
    L.in.c[c] = col_cur(a, (int) c);
    L.out.c[c] = (st.writes >> c) & 1u ? col_other(a, (int) c) : nullptr;
  }
  NListArgs nl_step{};
  if (st.near_agent >= 0) {
    AgentStore &b = sim->types[st.near_agent];
    const double r = step_radius(sim, st);
    const double rl = r * (1.0 + sim->skin);
    const bool want_list = list_eligible(sim, st, a);
    const bool gated = want_list && b.nl_ok && reach_for(b, rl) <= 4;
    if (!gated && !b.grid_valid) {
      int sp = span_open(sim, 2);
      int rc = build_grid(sim, b);
      span_close(sim, sp);
      if (rc) return rc;
      if (st.near_agent == st.agent) {
        for (size_t c = 0; c < a.cols.size(); c++) {
          L.in.c[c] = col_cur(a, (int) c);
          L.out.c[c] = (st.writes >> c) & 1u ? col_other(a, (int) c) : nullptr;
        }
      }
    }
    L.nb_n = (int) b.n;
    for (size_t c = 0; c < b.cols.size(); c++) L.nb.c[c] = col_cur(b, (int) c);
    L.nb_id = b.id[b.id_cur];
    L.grid = b.gv;
    L.grid.reach = reach_for(b, r);
    L.near_t = sqrtf_le_threshold(r);
    L.tile = sim->tile;
    L.lds_slack = sim->lds_slack;
    L.occupancy = b.ncells ? (float) ((double) b.n / (double) b.ncells) : 0.f;
    L.rc = (float) (r / b.cell * (1.0 + 1.0 / 1048576.0));
    // lists need the list radius inside the largest tile halo (4 cells)
    if (want_list && reach_for(b, rl) <= 4) {
      L.grid.reach = reach_for(b, rl);
      int sp = span_open(sim, 2);
      int rc = gated ? gated_rebuild(sim, b, st, L, r, rl) : enter_lists(sim, b, st, L, r, rl);
      span_close(sim, sp);
      if (rc) return rc;
      if (gated) {
        rc = list_rate(sim, b);
        if (rc) return rc;
      }
      nl_step = list_args(sim, b, r, rl);
      L.nl_mode = 2;
      L.nl = &nl_step;
    } else {
      b.nl_ok = false;
    }
    if (sim->order == 1) {
      int side = 2 * L.grid.reach + 1;
      int runs = side * side * (b.dim == 3 ? side : 1);
      if (runs > 27) return fail(ABL_ERR_UNSUPPORTED, "hip.order=index needs a stencil of at most 27 cells");
    }
  }
  L.self_id = a.id[a.id_cur];
  L.seed = sim->seed;
  L.timestep = sim->timestep;
  L.step_ordinal = ordinal;
  const int n_step = L.n;
  if (n_step > 0 && (st.removes || st.adds)) {
    L.dead = sim->dead;
    L.born = sim->born;
    hipLaunchKernelGGL(k_zero_u8, grid_blocks(n_step, 256), 256, 0, sim->stream, sim->dead, n_step);
    hipLaunchKernelGGL(k_zero_u8, grid_blocks(n_step, 256), 256, 0, sim->stream, sim->born, n_step);
    if (st.adds) {
      AgentStore &c = sim->types[st.adds - 1];
      for (size_t k = 0; k < c.cols.size(); k++) L.birth.c[k] = c.bmem[k];
    }
  }
  if (L.n > 0) {
    int sp = span_open(sim, 1);
    st.launch(L);
    span_close(sim, sp);
    HIP_TRY(hipGetLastError());
  }
  uint32_t w = st.writes;
  for (size_t c = 0; c < a.cols.size(); c++) {
    if ((w >> c) & 1u) a.cur[c] ^= 1;
  }
  if (a.pos_col >= 0) {
    for (int d = 0; d < a.dim; d++) {
      if ((w >> (a.pos_col + d)) & 1u) a.grid_valid = false;
    }
  }
  // Births first (their flags are indexed by the pre-step slots), then the
  // order-keeping removal; newborns of the same type sit after the parents
  // and are never removed, so the result is survivors in order followed by
  // newborns in parent-id order.
  int extra = 0;
  if (n_step > 0 && st.adds) {
    AgentStore &c = sim->types[st.adds - 1];
    const size_t before = c.n;
    int rc = apply_births(sim, a, c, n_step);
    if (rc) return rc;
    if (&c == &a) extra = (int) (c.n - before);
  }
  if (n_step > 0 && st.removes) {
    if (extra > 0) {
      hipLaunchKernelGGL(k_zero_u8, grid_blocks(extra, 256), 256, 0, sim->stream,
                         sim->dead + n_step, extra);
    }
    int rc = apply_removals(sim, a, n_step + extra);
    if (rc) return rc;
  }
  return ABL_OK;
}

extern "C" int abl_hip_simulate(abl_sim *sim, int timesteps) {
  CHECK_SIM(sim);
  if (timesteps < 0) return fail(ABL_ERR_ARG, "negative timestep count");
  sim->evused = 0;
  sim->spans.clear();
  if (sim->timing) HIP_TRY(hipEventRecord(sim->ev[0], sim->stream));
  auto t_prev = std::chrono::steady_clock::now();
  for (int t = 0; t < timesteps; t++) {
    uint32_t ordinal = 0;
    for (const StepDesc &st : sim->prog->steps) {
      int rc = launch_step(sim, st, ordinal++);
      if (rc) return rc;
    }
    if (sim->prog->seq_step || sim->prog->seq_step_prec) {
      // getLastExecTime(): seconds taken by this timestep's parallel steps
      HIP_TRY(hipStreamSynchronize(sim->stream));
      auto now = std::chrono::steady_clock::now();
      sim->last_exec = std::chrono::duration<double>(now - t_prev).count();
      int rc = sim->prog->seq_step_prec ? sim->prog->seq_step_prec(sim, sim->precision)
                                        : sim->prog->seq_step(sim, sim->params.data());
      if (rc) return rc;
      t_prev = std::chrono::steady_clock::now();
    }
    sim->timestep++;
  }
  if (sim->timing) {
    HIP_TRY(hipEventRecord(sim->ev[1], sim->stream));
    HIP_TRY(hipEventSynchronize(sim->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, sim->ev[0], sim->ev[1]));
    sim->last_ms[0] = ms;
    sim->last_ms[1] = sim->last_ms[2] = 0.0;
    for (auto &sp : sim->spans) {
      float m = 0.f;
      HIP_TRY(hipEventElapsedTime(&m, sp.second.first, sp.second.second));
      sim->last_ms[sp.first] += m;
    }
  }
  return ABL_OK;
}

extern "C" int abl_hip_synchronize(abl_sim *sim) {
  CHECK_SIM(sim);
  HIP_TRY(hipStreamSynchronize(sim->stream));
  return check_rebuild(sim);
}

extern "C" int abl_hip_query(abl_sim *sim, int *done) {
  CHECK_SIM(sim);
  if (!done) return fail(ABL_ERR_ARG, "null argument");
  hipError_t e = hipStreamQuery(sim->stream);
  if (e == hipSuccess) {
    *done = 1;
    return ABL_OK;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return ABL_OK;
  }
  return fail(ABL_ERR_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

// Device bytes a simulation of `program` with `agents` agents in total would
// hold (columns, ids, grid scratch, staging), for admission control.
extern "C" int abl_hip_estimate_footprint(const char *program, int precision, int num_params,
                                          const char *const *param_names,
                                          const double *param_values, size_t agents,
                                          size_t cells, size_t *bytes) {
  if (!program || !bytes || num_params < 0 || (num_params && (!param_names || !param_values)))
    return fail(ABL_ERR_ARG, "null argument");
  const ProgramDesc *prog = find_program(program);
  if (!prog) return fail(ABL_ERR_ARG, std::string("unknown program ") + program);
  std::vector<double> params;
  for (const ParamDesc &p : prog->params) params.push_back(p.value);
  for (int k = 0; k < num_params; k++) {
    size_t p = 0;
    while (p < prog->params.size() && std::strcmp(prog->params[p].name, param_names[k]) != 0) p++;
    if (p == prog->params.size())
      return fail(ABL_ERR_ARG, std::string("Unknown parameter \"") + param_names[k] + "\"");
    params[p] = param_values[k];
  }
  double lo[3], hi[3], gran = 0.0;
  prog->env(params.data(), lo, hi, &gran);
  bool adds = false, removes = false;
  for (const StepDesc &st : prog->steps) {
    adds = adds || st.adds;
    removes = removes || st.removes;
  }
  size_t b = 1u << 20;  // events, gate blocks, allocator rounding
  for (size_t t = 0; t < prog->agents.size(); t++) {
    AgentStore s;
    layout_agent(s, prog->agents[t], precision);
    const size_t cb = col_bytes(s);
    // columns, ids, sort scratch, upload/download staging
    size_t per_agent = 2 * cb + 8 + 32 + s.aos_size;
    if (adds) per_agent += 2 * cb;      // a parent's newborn staging, the child store's growth
    if (adds || removes) per_agent += 2 * 14;  // flags, their scan, birth lists (2 per agent)
    // the grid of this type, at the finest cell size a near loop over it uses
    size_t ncells = cells;
    double list_entries = 0.0;
    for (const StepDesc &st : prog->steps) {
      if (st.near_agent != (int) t || s.dim == 0 || !(gran > 0.0)) continue;
      double r = st.radius_fn ? st.radius_fn(precision) : 0.0;
      for (size_t p = 0; !st.radius_fn && p < prog->params.size(); p++)
        if (std::strcmp(prog->params[p].name, st.radius_param) == 0) r = st.radius_scale * params[p];
      const double cell = gran / 4.0;  // hip.cell_scale at most 4
      double nc = 1.0, vol = 1.0;
      for (int d = 0; d < s.dim; d++) {
        nc *= std::ceil((hi[d] - lo[d]) / cell) + 1.0;
        vol *= std::max(hi[d] - lo[d], cell);
      }
      ncells = std::max(ncells, (size_t) std::min(nc, 4e9));
      if (st.nlist && st.near_agent == st.agent && !st.removes && !st.adds) {
        // the list capacity enter_lists can reach at the default skin: the
        // estimate of a uniform population grown at most four times
        const double rl = r * (1.0 + kDefaultSkin);
        const double ball = s.dim == 3 ? 4.0 / 3.0 * M_PI * rl * rl * rl : M_PI * rl * rl;
        const double want = 1.5 * ball * (double) agents / vol + 16.0;
        list_entries = std::max(list_entries, std::min(4.0 * want + 8.0, 65528.0));
      }
    }
    if (list_entries > 0.0) per_agent += (size_t) (2.0 * list_entries) + kListMetaBytes;
    b += agents * per_agent + (ncells + 1) * (4 + 8 + 4 + 8 + 4);
  }
  *bytes = b;
  return ABL_OK;
}

extern "C" int abl_hip_footprint(abl_sim *sim, size_t *bytes) {
  CHECK_SIM(sim);
  if (!bytes) return fail(ABL_ERR_ARG, "null argument");
  size_t b = sim->staging_cap;
  for (const AgentStore &s : sim->types) b += store_bytes(s);
  if (sim->dead) b += sim->fcap * (1 + 1 + 4 + 4 + 4) + 4 + (sim->fcap / 2048 + 1) * 4;
  *bytes = b;
  return ABL_OK;
}

extern "C" int abl_hip_last_timing(abl_sim *sim, double out_ms[3]) {
  CHECK_SIM(sim);
  if (!out_ms) return fail(ABL_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) out_ms[k] = sim->last_ms[k];
  return ABL_OK;
}

extern "C" int abl_hip_neighbours(abl_sim *sim, int step, int64_t *pairs, size_t capacity,
                                  size_t *count) {
  CHECK_SIM(sim);
  if (step < 0 || step >= (int) sim->prog->steps.size()) return fail(ABL_ERR_ARG, "step out of range");
  const StepDesc &st = sim->prog->steps[step];
  if (count) *count = 0;
  if (st.near_agent < 0) return ABL_OK;
  AgentStore &a = sim->types[st.agent];
  AgentStore &b = sim->types[st.near_agent];
  if (!b.grid_valid) {
    int rc = build_grid(sim, b);
    if (rc) return rc;
  }
  const int n = (int) a.n;
  if (n == 0) return ABL_OK;
  const double r = step_radius(sim, st);
  GridView g = b.gv;
  g.reach = reach_for(b, r);
  const float t = sqrtf_le_threshold(r);
  uint32_t *cnt = nullptr, *off = nullptr;
  int64_t *dp = nullptr;
  ScopedDevice scratch;
  HIP_TRY(scratch.alloc(&cnt, (size_t) n * 4));
  HIP_TRY(scratch.alloc(&off, (size_t) n * 4));
  const int pa = a.pos_col, pb = b.pos_col;
  auto launch = [&](uint32_t *counts, const uint32_t *offsets, int64_t *out) {
    if (a.cols[pa].elem == 8) {
      hipLaunchKernelGGL(k_probe<double>, grid_blocks(n, 256), 256, 0, sim->stream,
                         (const double *) col_cur(a, pa), (const double *) col_cur(a, pa + 1),
                         (const double *) (a.dim == 3 ? col_cur(a, pa + 2) : nullptr), a.id[a.id_cur], n,
                         (const double *) col_cur(b, pb), (const double *) col_cur(b, pb + 1),
                         (const double *) (b.dim == 3 ? col_cur(b, pb + 2) : nullptr), b.id[b.id_cur], g,
                         t, counts, offsets, out);
    } else {
      hipLaunchKernelGGL(k_probe<float>, grid_blocks(n, 256), 256, 0, sim->stream,
                         (const float *) col_cur(a, pa), (const float *) col_cur(a, pa + 1),
                         (const float *) (a.dim == 3 ? col_cur(a, pa + 2) : nullptr), a.id[a.id_cur], n,
                         (const float *) col_cur(b, pb), (const float *) col_cur(b, pb + 1),
                         (const float *) (b.dim == 3 ? col_cur(b, pb + 2) : nullptr), b.id[b.id_cur], g,
                         t, counts, offsets, out);
    }
  };
  launch(cnt, nullptr, nullptr);
  std::vector<uint32_t> hc((size_t) n);
  HIP_TRY(hipMemcpyAsync(hc.data(), cnt, (size_t) n * 4, hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(hipStreamSynchronize(sim->stream));
  std::vector<uint32_t> ho((size_t) n);
  size_t total = 0;
  for (int i = 0; i < n; i++) {
    ho[i] = (uint32_t) total;
    total += hc[i];
  }
  if (count) *count = total;
  if (pairs && total > 0) {
    HIP_TRY(hipMemcpyAsync(off, ho.data(), (size_t) n * 4, hipMemcpyHostToDevice, sim->stream));
    HIP_TRY(scratch.alloc(&dp, total * 16));
    launch(nullptr, off, dp);
    std::vector<int64_t> hp(total * 2);
    HIP_TRY(hipMemcpyAsync(hp.data(), dp, total * 16, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    std::vector<std::pair<int64_t, int64_t>> v(total);
    for (size_t k = 0; k < total; k++) v[k] = {hp[2 * k], hp[2 * k + 1]};
    std::sort(v.begin(), v.end());
    size_t m = std::min(total, capacity);
    for (size_t k = 0; k < m; k++) {
      pairs[2 * k] = v[k].first;
      pairs[2 * k + 1] = v[k].second;
    }
  }
  HIP_TRY(hipGetLastError());
  return ABL_OK;
}

// ---- sequential-step builtins (src/main.cpp:66-130) -------------------------------

static int member_columns(abl_sim *sim, int t, const char *member, int *first, int *ncomp) {
  if (t < 0 || t >= (int) sim->types.size() || !member) return fail(ABL_ERR_ARG, "bad argument");
  const AgentStore &s = sim->types[t];
  for (size_t c = 0; c < s.cols.size(); c++) {
    if (std::strcmp(s.desc->members[s.cols[c].member].name, member) == 0) {
      *first = (int) c;
      *ncomp = kind_components(s.desc->members[s.cols[c].member].kind);
      return ABL_OK;
    }
  }
  return fail(ABL_ERR_ARG, std::string("Agent has no member \"") + member + "\"");
}

// sum(Type.member): bool sums count the true values; vectors sum per component.
extern "C" int abl_hip_reduce_sum(abl_sim *sim, int t, const char *member, double *out) {
  CHECK_SIM(sim);
  int c0 = 0, nc = 0;
  int rc = member_columns(sim, t, member, &c0, &nc);
  if (rc) return rc;
  if (!out) return fail(ABL_ERR_ARG, "null output");
  AgentStore &s = sim->types[t];
  const int n = (int) s.n;
  const int blocks = std::max(1, std::min(1024, (n + 255) / 256));
  const int kind = s.desc->members[s.cols[c0].member].kind;
  void *partial = nullptr;
  ScopedDevice scratch;
  HIP_TRY(scratch.alloc(&partial, (size_t) blocks * 8));
  for (int k = 0; k < nc; k++) {
    const void *col = col_cur(s, c0 + k);
    if (n == 0) {
      out[k] = 0.0;
      continue;
    }
    if (kind == K_BOOL) {
      hipLaunchKernelGGL((k_reduce_sum<uint8_t, long long>), blocks, 256, 0, sim->stream,
                         (const uint8_t *) col, n, (long long *) partial);
    } else if (kind == K_INT) {
      hipLaunchKernelGGL((k_reduce_sum<int32_t, long long>), blocks, 256, 0, sim->stream,
                         (const int32_t *) col, n, (long long *) partial);
    } else if (s.cols[c0].elem == 8) {
      hipLaunchKernelGGL((k_reduce_sum<double, double>), blocks, 256, 0, sim->stream,
                         (const double *) col, n, (double *) partial);
    } else {
      hipLaunchKernelGGL((k_reduce_sum<float, double>), blocks, 256, 0, sim->stream,
                         (const float *) col, n, (double *) partial);
    }
    HIP_TRY(hipGetLastError());
    std::vector<char> h((size_t) blocks * 8);
    HIP_TRY(hipMemcpyAsync(h.data(), partial, h.size(), hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    double acc = 0.0;
    long long iacc = 0;
    for (int b = 0; b < blocks; b++) {
      if (kind == K_BOOL || kind == K_INT) iacc += ((long long *) h.data())[b];
      else acc += ((double *) h.data())[b];
    }
    out[k] = (kind == K_BOOL || kind == K_INT) ? (double) iacc : acc;
  }
  return ABL_OK;
}

// ---- sum(Type.member) without allocation ----------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that forbid it): the per-block partial sums live in
// a file-scope device array, one copy per GPU, whose address is looked up
// once per device.  Host code never passes the array's name to a kernel or
// copy, only that address.  The kernel writes every partial it launches
// before the host reads them back, so nothing needs clearing between calls;
// calls on one device are serialised by a host lock, since they share the
// array.

constexpr int kNoAllocPartials = 1024;  // blocks of the reduction, at most
static __device__ __align__(256) double scratch[kNoAllocPartials];  // 8 KiB per GPU
constexpr int kMaxDevices = 64;

static std::mutex g_scratch_mu[kMaxDevices];
static void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `dev` (looked up once, with dev current).
static int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(ABL_ERR_UNSUPPORTED, "device index out of range");
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t e2 = hipSetDevice(prev);
    if (e != hipSuccess)
      return fail(ABL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(ABL_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e2));
    if (!p) return fail(ABL_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ABL_OK;
}

extern "C" int abl_hip_reduce_sum_noalloc(abl_sim *sim, int t, const char *member, double *out) {
  CHECK_SIM(sim);
  int c0 = 0, nc = 0;
  int rc = member_columns(sim, t, member, &c0, &nc);
  if (rc) return rc;
  if (!out) return fail(ABL_ERR_ARG, "null output");
  AgentStore &s = sim->types[t];
  const int n = (int) s.n;
  // every agent count fits: the kernel strides over the column with at most
  // kNoAllocPartials blocks, one partial each
  const int blocks = std::max(1, std::min(kNoAllocPartials, (n + 255) / 256));
  if (blocks > kNoAllocPartials) return fail(ABL_ERR_ARG, "reduction exceeds the scratch array");
  const int kind = s.desc->members[s.cols[c0].member].kind;
  std::lock_guard<std::mutex> lock(g_scratch_mu[sim->device < 0 || sim->device >= kMaxDevices ? 0 : sim->device]);
  void *ptr = nullptr;
  rc = scratch_address(sim->device, &ptr);
  if (rc) return rc;
  double partial[kNoAllocPartials];  // host copy of the partials (stack, no allocation)
  for (int k = 0; k < nc; k++) {
    const void *col = col_cur(s, c0 + k);
    if (n == 0) {
      out[k] = 0.0;
      continue;
    }
    if (kind == K_BOOL) {
      hipLaunchKernelGGL((k_reduce_sum<uint8_t, long long>), blocks, 256, 0, sim->stream,
                         (const uint8_t *) col, n, (long long *) ptr);
    } else if (kind == K_INT) {
      hipLaunchKernelGGL((k_reduce_sum<int32_t, long long>), blocks, 256, 0, sim->stream,
                         (const int32_t *) col, n, (long long *) ptr);
    } else if (s.cols[c0].elem == 8) {
      hipLaunchKernelGGL((k_reduce_sum<double, double>), blocks, 256, 0, sim->stream,
                         (const double *) col, n, (double *) ptr);
    } else {
      hipLaunchKernelGGL((k_reduce_sum<float, double>), blocks, 256, 0, sim->stream,
                         (const float *) col, n, (double *) ptr);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(partial, ptr, (size_t) blocks * 8, hipMemcpyDeviceToHost, sim->stream));
    HIP_TRY(hipStreamSynchronize(sim->stream));
    double acc = 0.0;
    long long iacc = 0;
    for (int b = 0; b < blocks; b++) {
      if (kind == K_BOOL || kind == K_INT) iacc += ((const long long *) partial)[b];
      else acc += partial[b];
    }
    out[k] = (kind == K_BOOL || kind == K_INT) ? (double) iacc : acc;
  }
  return ABL_OK;
}

// count(Type.member, value): agents whose (scalar) member equals value.
extern "C" int abl_hip_reduce_count_eq(abl_sim *sim, int t, const char *member, double value,
                                       size_t *out) {
  CHECK_SIM(sim);
  int c0 = 0, nc = 0;
  int rc = member_columns(sim, t, member, &c0, &nc);
  if (rc) return rc;
  if (!out) return fail(ABL_ERR_ARG, "null output");
  if (nc != 1) return fail(ABL_ERR_UNSUPPORTED, "count(member, value) on a vector member");
  AgentStore &s = sim->types[t];
  const int n = (int) s.n;
  *out = 0;
  if (n == 0) return ABL_OK;
  const int blocks = std::max(1, std::min(1024, (n + 255) / 256));
  const int kind = s.desc->members[s.cols[c0].member].kind;
  unsigned long long *partial = nullptr;
  ScopedDevice scratch;
  HIP_TRY(scratch.alloc(&partial, (size_t) blocks * 8));
  const void *col = col_cur(s, c0);
  if (kind == K_BOOL) {
    hipLaunchKernelGGL(k_reduce_count_eq<uint8_t>, blocks, 256, 0, sim->stream,
                       (const uint8_t *) col, n, (uint8_t) (value != 0.0), partial);
  } else if (kind == K_INT) {
    hipLaunchKernelGGL(k_reduce_count_eq<int32_t>, blocks, 256, 0, sim->stream,
                       (const int32_t *) col, n, (int32_t) value, partial);
  } else if (s.cols[c0].elem == 8) {
    hipLaunchKernelGGL(k_reduce_count_eq<double>, blocks, 256, 0, sim->stream,
                       (const double *) col, n, value, partial);
  } else {
    hipLaunchKernelGGL(k_reduce_count_eq<float>, blocks, 256, 0, sim->stream,
                       (const float *) col, n, (float) value, partial);
  }
  HIP_TRY(hipGetLastError());
  std::vector<unsigned long long> h((size_t) blocks);
  HIP_TRY(hipMemcpyAsync(h.data(), partial, h.size() * 8, hipMemcpyDeviceToHost, sim->stream));
  HIP_TRY(hipStreamSynchronize(sim->stream));
  unsigned long long acc = 0;
  for (unsigned long long v : h) acc += v;
  *out = (size_t) acc;
  return ABL_OK;
}

// getLastExecTime(): wall-clock seconds of the last timestep's parallel steps.
extern "C" int abl_hip_last_exec_time(abl_sim *sim, double *seconds) {
  CHECK_SIM(sim);
  if (!seconds) return fail(ABL_ERR_ARG, "null output");
  *seconds = sim->last_exec;
  return ABL_OK;
}

// log_csv(...): append one row; the runtime keeps the text and, when
// "hip.log_csv" names a file, appends it there too.
extern "C" int abl_hip_log_csv(abl_sim *sim, const char *row) {
  CHECK_SIM(sim);
  if (!row) return fail(ABL_ERR_ARG, "null row");
  sim->log += row;
  if (!sim->log_path.empty()) {
    FILE *f = std::fopen(sim->log_path.c_str(), "a");
    if (!f) return fail(ABL_ERR_ARG, "cannot open " + sim->log_path);
    std::fputs(row, f);
    std::fclose(f);
  }
  return ABL_OK;
}

extern "C" int abl_hip_log_text(abl_sim *sim, char *buf, size_t capacity, size_t *len) {
  CHECK_SIM(sim);
  if (len) *len = sim->log.size();
  if (buf && capacity) {
    size_t m = std::min(capacity - 1, sim->log.size());
    std::memcpy(buf, sim->log.data(), m);
    buf[m] = 0;
  }
  return ABL_OK;
}
