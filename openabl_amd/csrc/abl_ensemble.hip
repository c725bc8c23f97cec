// abl_ensemble.hip -- many independent runs of one program batched into one
// launch per kernel, with one timestep captured as a HIP graph.
//
// Members (uneven agent counts, per-member parameters such as box size and
// flocking weights) are stored back to back in one set of columns; member m
// owns slots [off[m], off[m+1]) and the grid cells [cell_base[m],
// cell_base[m+1]), laid out member-major, so one counting sort over all cells
// keeps every member cell-sorted inside its own slot range.  The neighbour
// kernel is the program's LDS-tiled kernel over the concatenated tile list
// (tiled_step_multi); each block finds its member and reads that member's
// arguments through the scalar cache.  A member's agents are ordered, and its
// neighbours visited, exactly as in a run of it alone, so each member's result
// is bit-identical to that run.
//
// Neighbour lists ("hip.nlist", default on when the members share the near
// radius): one set of lists over the concatenated slots, built after a sort
// and walked while no agent of the member has moved more than skin / 2 since
// (abl_tile.hpp).  Each member has its own rebuild gate, set by its own
// agents, so it rebuilds on exactly the timesteps a run of it alone would.
// The sort runs when some member is due; a member that is not keeps its
// stored cell keys, so the sort leaves its agents where they are.  The sort
// gathers into the other buffers and the list build copies the agents of the
// rebuilt members back, so the step always reads the current buffers.
//
// With static populations every timestep issues the same kernels on the same
// buffers (without lists the grid gather and the step each flip the columns;
// with lists the step flips them, and two timesteps return to the start): two
// timesteps are captured once as a graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/abl_hip.h"
#include "abl_pipeline.hpp"
#include "abl_runtime.hpp"

using namespace abl;

namespace {

thread_local std::string g_ens_error;

int ens_fail(int code, const std::string &msg) {
  g_ens_error = msg;
  return code;
}

#define ENS_TRY(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      return ens_fail(ABL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                              \
  } while (0)

struct Member {
  int n = 0;
  uint32_t off = 0;
  std::vector<double> params;
  GridView g{};
  uint32_t cell_base = 0, tile_base = 0, ntiles = 0;
  double cell = 0.0;
  float near_t = 0.f, rc = 0.f;
};

// member of slot i (binary search over the slot offsets)
__device__ __forceinline__ int member_of(const uint32_t *off, int members, uint32_t i) {
  int lo = 0, hi = members - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Every slot's cell key.  With a gate: only while *gate is set, and a member
// whose own gate (mgate[m]) is clear keeps the key of the slot's cell at the
// last sort (skey), so the sort leaves its agents in place.
template <typename P>
__global__ void k_ens_bin(const P *__restrict__ px, const P *__restrict__ py,
                          const uint32_t *__restrict__ off, const GridView *__restrict__ grids,
                          const uint32_t *__restrict__ cell_base, int members,
                          uint32_t *__restrict__ key, uint32_t *__restrict__ count, int n,
                          const uint32_t *__restrict__ gate = nullptr,
                          const uint32_t *__restrict__ mgate = nullptr,
                          const uint32_t *__restrict__ skey = nullptr) {
  if (gate && *gate == 0u) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = member_of(off, members, (uint32_t) i);
  uint32_t k;
  if (mgate && mgate[m] == 0u) {
    k = skey[i];
  } else {
    const GridView &g = grids[m];
    const int3 c = cell_of<P>(g, px[i], py[i], (P) 0);
    k = cell_base[m] + (uint32_t) c.y * (uint32_t) g.n[0] + (uint32_t) c.x;
  }
  key[i] = k;
  atomicAdd(&count[k], 1u);
}

// slot s of member m back to record id within that member's output block
__global__ void k_ens_unpermute(char *__restrict__ aos, const void *const *__restrict__ cols,
                                int elem, int ncols, const int32_t *__restrict__ id, uint32_t begin,
                                int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = begin + (uint32_t) k;
  char *rec = aos + (size_t) id[s] * (size_t) (elem * ncols);
  for (int c = 0; c < ncols; c++) {
    if (elem == 8) ((double *) rec)[c] = ((const double *) cols[c])[s];
    else ((float *) rec)[c] = ((const float *) cols[c])[s];
  }
}

}  // namespace

struct abl_ensemble {
  const ProgramDesc *prog = nullptr;
  const EnsembleHooks *hooks = nullptr;
  int precision = 0;
  std::vector<Member> mem;
  size_t n = 0;
  int ncols = 0, elem = 8;
  void *col[2][kMaxCols] = {};
  int cur = 0;
  int32_t *id[2] = {nullptr, nullptr};
  int id_cur = 0;
  uint32_t *d_off = nullptr, *d_cell_base = nullptr, *d_tile_base = nullptr;
  GridView *d_grids = nullptr;
  void *d_args[2] = {nullptr, nullptr};  // member arguments for cur = 0 / cur = 1
  uint32_t *key = nullptr, *count = nullptr, *start = nullptr, *bsum = nullptr, *tmp_idx = nullptr;
  int32_t *tmp_id = nullptr;
  size_t ncells = 0;
  unsigned total_tiles = 0;
  int reach = 1;
  float occupancy = 0.f;
  int cell_scale = 2, tile = 1, block = 256;
  bool use_graph = true;
  bool ready = false;
  uint64_t timestep = 0;       // timesteps issued (the lists' gates go by its parity)
  // neighbour lists
  bool nlist = true;           // "hip.nlist": lists when the members share the near radius
  double skin = kDefaultSkin;  // "hip.skin": list skin as a fraction of the near radius
  int list_cap = 0;            // "hip.list_cap": entries per list (0: sized from the data)
  bool lists = false;          // in use (decided by abl_ens_init)
  double r = 0.0, rl = 0.0, cell = 0.0;
  int nl_cap = 0;
  uint16_t *nl_idx = nullptr, *nl_len = nullptr;
  uint32_t *nl_perm = nullptr;
  uint4 *nl_prow0 = nullptr;
  double *nl_pos0 = nullptr;
  uint32_t *nl_gate = nullptr;   // [0..1] some member due, by parity; [2] longest list
  uint32_t *nl_mgate = nullptr;  // [2][members] member due, by parity
  uint32_t *skey = nullptr;      // each slot's cell key at the last sort
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
};

extern "C" const char *abl_ens_last_error(void) { return g_ens_error.c_str(); }

// Member m's generator start state: the base state with its first word mixed
// with m, so members draw independent, reproducible streams.
extern "C" int abl_ens_member_seed(uint64_t s0, uint64_t s1, int member, uint64_t out[2]) {
  if (!out || member < 0) return ens_fail(ABL_ERR_ARG, "bad argument");
  out[0] = s0 ^ (0x9E3779B97F4A7C15ULL * (uint64_t) (member + 1));
  out[1] = s1;
  if (out[0] == 0 && out[1] == 0) out[1] = 1;
  return ABL_OK;
}

extern "C" int abl_ens_create(const char *program, int precision, int members,
                              const int *counts, int num_params, const char *const *names,
                              const double *values, abl_ensemble **out) {
  if (!program || !out || members < 1 || !counts || (num_params && (!names || !values)))
    return ens_fail(ABL_ERR_ARG, "bad argument");
  *out = nullptr;
  const ProgramDesc *prog = find_program(program);
  if (!prog) return ens_fail(ABL_ERR_ARG, std::string("unknown program ") + program);
  if (!prog->ensemble) return ens_fail(ABL_ERR_UNSUPPORTED, "program has no ensemble support");
  if (precision != ABL_PREC_F64 && precision != ABL_PREC_F32)
    return ens_fail(ABL_ERR_UNSUPPORTED, "precision not supported");
  abl_ensemble *e = new abl_ensemble();
  e->prog = prog;
  e->hooks = prog->ensemble;
  e->precision = precision;
  e->mem.resize((size_t) members);
  for (int m = 0; m < members; m++) {
    Member &mb = e->mem[(size_t) m];
    for (const ParamDesc &p : prog->params) mb.params.push_back(p.value);
    for (int k = 0; k < num_params; k++) {
      bool found = false;
      for (size_t p = 0; p < prog->params.size(); p++) {
        if (std::strcmp(prog->params[p].name, names[k]) == 0) {
          mb.params[p] = values[(size_t) m * (size_t) num_params + (size_t) k];
          found = true;
        }
      }
      if (!found) {
        delete e;
        return ens_fail(ABL_ERR_ARG, std::string("Unknown parameter \"") + names[k] + "\"");
      }
    }
    mb.n = counts[m];
  }
  e->elem = precision == ABL_PREC_F64 ? 8 : 4;
  int ncols = 0;
  for (const MemberDesc &md : prog->agents[0].members) ncols += kind_components(md.kind);
  e->ncols = ncols;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
    delete e;
    return ens_fail(ABL_ERR_DEVICE, "stream/event creation failed");
  }
  *out = e;
  return ABL_OK;
}

extern "C" int abl_ens_set_config(abl_ensemble *e, const char *key, const char *value) {
  if (!e || !key || !value) return ens_fail(ABL_ERR_ARG, "bad argument");
  if (e->ready) return ens_fail(ABL_ERR_STATE, "configure before abl_ens_init");
  std::string k(key), v(value);
  if (k == "hip.cell_scale") {
    e->cell_scale = std::max(1, std::min(4, std::atoi(value)));
  } else if (k == "hip.tile") {
    e->tile = std::max(1, std::min(3, std::atoi(value)));
  } else if (k == "hip.block") {
    int b = std::atoi(value);
    if (b < 64 || b > 256 || b % 64) return ens_fail(ABL_ERR_ARG, "hip.block must be 64..256");
    e->block = b;
  } else if (k == "hip.graph") {
    e->use_graph = v == "1" || v == "true";
  } else if (k == "hip.nlist") {
    if (v != "0" && v != "1") return ens_fail(ABL_ERR_ARG, "hip.nlist must be 0 or 1");
    e->nlist = v == "1";
  } else if (k == "hip.skin") {
    const double f = std::atof(value);
    if (!(f > 0.0 && f <= 1.0)) return ens_fail(ABL_ERR_ARG, "hip.skin must be in (0, 1]");
    e->skin = f;
  } else if (k == "hip.list_cap") {
    const int c = std::atoi(value);
    if (c < 0 || c > 65528) return ens_fail(ABL_ERR_ARG, "hip.list_cap must be 0..65528");
    e->list_cap = c;
  } else {
    return ens_fail(ABL_ERR_ARG, "unknown config key " + k);
  }
  return ABL_OK;
}

static void free_all(abl_ensemble *e) {
  for (int b = 0; b < 2; b++) {
    for (int c = 0; c < kMaxCols; c++) {
      if (e->col[b][c]) (void) hipFree(e->col[b][c]);
      e->col[b][c] = nullptr;
    }
    if (e->id[b]) (void) hipFree(e->id[b]);
    if (e->d_args[b]) (void) hipFree(e->d_args[b]);
    e->id[b] = nullptr;
    e->d_args[b] = nullptr;
  }
  void *p[] = {e->d_off, e->d_cell_base, e->d_tile_base, e->d_grids, e->key, e->count,
               e->start, e->bsum, e->tmp_idx, e->tmp_id, e->nl_idx, e->nl_len, e->nl_perm,
               e->nl_prow0, e->nl_pos0, e->nl_gate, e->nl_mgate, e->skey};
  for (void *q : p) {
    if (q) (void) hipFree(q);
  }
  e->nl_idx = e->nl_len = nullptr;
  e->nl_perm = e->nl_gate = e->nl_mgate = e->skey = nullptr;
  e->nl_prow0 = nullptr;
  e->nl_pos0 = nullptr;
  e->d_off = e->d_cell_base = e->d_tile_base = e->key = e->count = e->start = e->bsum = nullptr;
  e->tmp_idx = nullptr;
  e->tmp_id = nullptr;
  e->d_grids = nullptr;
  if (e->exec) (void) hipGraphExecDestroy(e->exec);
  if (e->graph) (void) hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

// The list arguments of the current timestep (as a run alone forms them).
static NListArgs ens_list_args(const abl_ensemble *e) {
  NListArgs nl{};
  nl.idx = e->nl_idx;
  nl.len = e->nl_len;
  nl.perm = e->nl_perm;
  nl.prow0 = e->nl_prow0;
  nl.pos0 = e->nl_pos0;
  nl.gate = e->nl_gate;
  nl.stats = e->nl_gate + 2;
  nl.n = (int) e->n;
  nl.cap = e->nl_cap;
  nl.stride = (int) e->n;
  nl.parity = (int) (e->timestep & 1u);
  nl.rc_list = (float) (e->rl / e->cell * (1.0 + 1.0 / 1048576.0));
  const double rw = e->rl * (1.0 + 1e-6) + 1e-5 * e->r;
  nl.r2_list = rw * rw;
  const double h = 0.5 * (e->rl - e->r);
  nl.half_skin2 = h * h;
  nl.ncells = (int) e->ncells;
  return nl;
}

// The gated rebuild of the current timestep: the sort of every slot into the
// other buffers (members not due keep their places), then the lists of the
// members due, which copy their agents back to the current buffers.
static void issue_rebuild(abl_ensemble *e) {
  const int n = (int) e->n;
  const int M = (int) e->mem.size();
  hipStream_t st = e->stream;
  const int b = e->cur;
  const uint32_t par = (uint32_t) (e->timestep & 1u);
  const uint32_t *gate = e->nl_gate + par;
  const uint32_t *mg = e->nl_mgate + (size_t) par * (size_t) M;
  if (e->elem == 8) {
    hipLaunchKernelGGL(k_ens_bin<double>, grid_blocks(n, 256), 256, 0, st,
                       (const double *) e->col[b][0], (const double *) e->col[b][1], e->d_off,
                       e->d_grids, e->d_cell_base, M, e->key, e->count, n, gate, mg, e->skey);
  } else {
    hipLaunchKernelGGL(k_ens_bin<float>, grid_blocks(n, 256), 256, 0, st,
                       (const float *) e->col[b][0], (const float *) e->col[b][1], e->d_off,
                       e->d_grids, e->d_cell_base, M, e->key, e->count, n, gate, mg, e->skey);
  }
  const int nc = (int) e->ncells;
  const int ntiles = (nc + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_scan_tiles<uint32_t, false>), ntiles, kScanBlock, 0, st, e->count, e->start,
                     e->bsum, nc, gate);
  hipLaunchKernelGGL(k_scan_totals, 1, 1024, 0, st, e->bsum, ntiles, e->start, nc, gate);
  hipLaunchKernelGGL(k_scan_add, grid_blocks(nc, 256), 256, 0, st, e->start, e->bsum, nc, gate);
  hipLaunchKernelGGL(k_scatter, grid_blocks(n, 256), 256, 0, st, e->key, e->count, e->start,
                     e->id[e->id_cur], e->tmp_idx, e->tmp_id, n, gate);
  GatherCols gc{};
  gc.ncols = e->ncols;
  for (int c = 0; c < e->ncols; c++) {
    gc.src[c] = e->col[b][c];
    gc.dst[c] = e->col[1 - b][c];
    gc.elem[c] = e->elem;
  }
  hipLaunchKernelGGL(k_rank_gather, grid_blocks(n, 256), 256, 0, st, e->tmp_idx, e->tmp_id,
                     e->key, e->start, gc, e->id[1 - e->id_cur], n, e->skey, nullptr, gate);
  NListArgs nl = ens_list_args(e);
  nl.gated = 1;
  nl.copy.ncols = e->ncols;
  for (int c = 0; c < e->ncols; c++) {
    nl.copy.src[c] = e->col[1 - b][c];
    nl.copy.dst[c] = e->col[b][c];
    nl.copy.elem[c] = e->elem;
  }
  nl.id_src = e->id[1 - e->id_cur];
  nl.id_dst = e->id[e->id_cur];
  nl.skey_src = e->skey;
  nl.skey_dst = e->skey;
  // the build reads the sorted buffers: the arguments whose input is 1 - b
  e->hooks->launch_multi_list(e->precision, e->d_args[b], e->d_tile_base, M, e->total_tiles,
                              e->reach, e->occupancy, e->tile, e->block, st, nl, e->nl_mgate, 1);
}

// Every member due a rebuild at this parity (p) and none at the other.
static int set_gates(abl_ensemble *e, uint32_t p) {
  const size_t M = e->mem.size();
  std::vector<uint32_t> g(16, 0u), mg(2 * M, 0u);
  g[p] = 1u;
  for (size_t m = 0; m < M; m++) mg[p * M + m] = 1u;
  ENS_TRY(hipMemcpy(e->nl_gate, g.data(), 16 * 4, hipMemcpyHostToDevice));
  ENS_TRY(hipMemcpy(e->nl_mgate, mg.data(), 2 * M * 4, hipMemcpyHostToDevice));
  return ABL_OK;
}

// List buffers, and every member's first sort and lists (as a run alone
// enters list mode: capacity "hip.list_cap", or from the expected population
// of the list radius, grown once to the longest list, at most four times the
// estimate).
static int init_lists(abl_ensemble *e) {
  const size_t N = e->n;
  const size_t M = e->mem.size();
  int want = 16;
  for (const Member &mb : e->mem) {
    const double env = (double) mb.g.n[0] * (double) mb.g.n[1] * mb.cell * mb.cell;
    const double expect = env > 0 ? M_PI * e->rl * e->rl * (double) mb.n / env : 0.0;
    want = std::max(want, (int) (1.5 * expect) + 16);
  }
  if (e->list_cap) want = e->list_cap;
  int cap = std::min((want + 7) & ~7, 65528);
  ENS_TRY(hipMalloc(&e->nl_len, N * 2));
  ENS_TRY(hipMalloc(&e->nl_perm, N * 4));
  ENS_TRY(hipMalloc(&e->nl_prow0, N * 16));
  ENS_TRY(hipMalloc(&e->nl_pos0, 2 * N * 8));
  ENS_TRY(hipMalloc(&e->nl_gate, 16 * 4));
  ENS_TRY(hipMalloc(&e->nl_mgate, 2 * M * 4));
  ENS_TRY(hipMalloc(&e->skey, N * 4));
  for (int attempt = 0;; attempt++) {
    if (e->nl_idx) ENS_TRY(hipFree(e->nl_idx));
    e->nl_idx = nullptr;
    ENS_TRY(hipMalloc(&e->nl_idx, (size_t) cap * N * 2));
    e->nl_cap = cap;
    int rc = set_gates(e, 0u);
    if (rc) return rc;
    issue_rebuild(e);
    ENS_TRY(hipGetLastError());
    ENS_TRY(hipStreamSynchronize(e->stream));
    uint32_t longest = 0;
    ENS_TRY(hipMemcpy(&longest, e->nl_gate + 2, 4, hipMemcpyDeviceToHost));
    const int grow = std::min((int) (1.25 * longest) + 16, 4 * want);
    if (attempt > 0 || (int) longest <= cap || grow <= cap || cap >= 65528 || e->list_cap) break;
    cap = std::min((grow + 7) & ~7, 65528);
  }
  // the lists are those of the initial state: no member is due at the first
  // timestep (an overflowed list sets its member's gate at that step)
  std::vector<uint32_t> zero(std::max<size_t>(16, 2 * M), 0u);
  ENS_TRY(hipMemcpy(e->nl_gate, zero.data(), 16 * 4, hipMemcpyHostToDevice));
  ENS_TRY(hipMemcpy(e->nl_mgate, zero.data(), 2 * M * 4, hipMemcpyHostToDevice));
  return ABL_OK;
}

// Initial state of every member (main()'s agent creation from the member's
// generator state), the grids, and the per-member kernel arguments.
extern "C" int abl_ens_init(abl_ensemble *e, uint64_t s0, uint64_t s1) {
  if (!e) return ens_fail(ABL_ERR_ARG, "bad argument");
  free_all(e);
  const int M = (int) e->mem.size();
  std::vector<char> host;
  size_t total = 0, cells = 0;
  unsigned tiles = 0;
  int tx = 16, ty = 8, tz = 1;
  e->hooks->tile_shape(e->tile, &tx, &ty, &tz);
  e->reach = 1;
  double occ_max = 0.0;
  // the near radius of member m (under use_float the parameter is a float
  // global, as in a run alone)
  auto radius_of = [&](const Member &mb) {
    const StepDesc &sd = e->prog->steps[0];
    double r = 0.0;
    for (size_t p = 0; p < e->prog->params.size(); p++) {
      if (std::strcmp(e->prog->params[p].name, sd.radius_param) == 0) {
        double v = mb.params[p];
        if (e->precision != ABL_PREC_F64) v = (double) (float) v;
        r = sd.radius_scale * v;
      }
    }
    return r;
  };
  // neighbour lists when every member has the same near radius and cell size
  // and the list radius needs no more than four cells of reach (as for a run
  // alone); their cells are widened by 1 + skin, as a run alone's are
  e->lists = e->nlist && e->hooks->launch_multi_list && M > 0;
  double gran0 = 0.0;
  for (int m = 0; m < M && e->lists; m++) {
    double lo[3], hi[3], gran;
    e->prog->env(e->mem[(size_t) m].params.data(), lo, hi, &gran);
    const double r = radius_of(e->mem[(size_t) m]);
    if (m == 0) {
      gran0 = gran;
      e->r = r;
    } else if (gran != gran0 || r != e->r) {
      e->lists = false;
    }
  }
  if (e->lists) {
    e->rl = e->r * (1.0 + e->skin);
    const double cell = gran0 / (double) e->cell_scale * (1.0 + 1.0 / 1048576.0) * (1.0 + e->skin);
    if (!(e->r > 0.0) || std::ceil(e->rl / cell * (1.0 + 1.0 / 2097152.0)) > 4.0) e->lists = false;
  }
  const double inflate = e->lists ? 1.0 + e->skin : 1.0;
  for (int m = 0; m < M; m++) {
    Member &mb = e->mem[(size_t) m];
    uint64_t st[2];
    abl_ens_member_seed(s0, s1, m, st);
    abl_rng_seed(st[0], st[1]);
    std::vector<char> recs;
    size_t cnt = 0;
    int rc = e->hooks->host_records(mb.params.data(), e->precision, recs, &cnt);
    if (rc) return ens_fail(rc, "host initialisation failed");
    mb.n = (int) cnt;
    mb.off = (uint32_t) total;
    total += cnt;
    host.insert(host.end(), recs.begin(), recs.end());
    double lo[3], hi[3], gran;
    e->prog->env(mb.params.data(), lo, hi, &gran);
    const double cell = gran / (double) e->cell_scale * (1.0 + 1.0 / 1048576.0) * inflate;
    GridView g{};
    g.dim = 2;
    for (int d = 0; d < 3; d++) {
      g.lo[d] = d < 2 ? lo[d] : 0.0;
      g.n[d] = d < 2 ? std::max(1, (int) std::ceil((hi[d] - lo[d]) / cell)) : 1;
    }
    g.inv_cell = 1.0 / cell;
    const double r = radius_of(mb);
    // the stencil reaches the list radius in list mode
    const double reach_r = e->lists ? e->rl : r;
    g.reach = std::max(1, (int) std::ceil(reach_r / cell * (1.0 + 1.0 / 2097152.0)));
    e->reach = std::max(e->reach, g.reach);
    mb.g = g;
    mb.cell = cell;
    mb.near_t = sqrtf_le_threshold(r);
    mb.rc = (float) (r / cell * (1.0 + 1.0 / 1048576.0));
    mb.cell_base = (uint32_t) cells;
    cells += (size_t) g.n[0] * (size_t) g.n[1];
    mb.tile_base = tiles;
    mb.ntiles = (unsigned) (((g.n[0] + tx - 1) / tx) * ((g.n[1] + ty - 1) / ty));
    tiles += mb.ntiles;
    occ_max = std::max(occ_max, (double) cnt / ((double) g.n[0] * g.n[1]));
  }
  for (Member &mb : e->mem) mb.g.reach = e->reach;  // one kernel instance for all members
  if (e->lists) e->cell = e->mem[0].cell;
  e->n = total;
  e->ncells = cells;
  e->total_tiles = tiles;
  e->occupancy = (float) occ_max;
  const size_t N = std::max<size_t>(total, 1);
  for (int b = 0; b < 2; b++) {
    for (int c = 0; c < e->ncols; c++) ENS_TRY(hipMalloc(&e->col[b][c], N * (size_t) e->elem));
    ENS_TRY(hipMalloc(&e->id[b], N * 4));
  }
  ENS_TRY(hipMalloc(&e->key, N * 4));
  ENS_TRY(hipMalloc(&e->tmp_idx, N * 4));
  ENS_TRY(hipMalloc(&e->tmp_id, N * 4));
  ENS_TRY(hipMalloc(&e->count, (cells + 1) * 4));
  ENS_TRY(hipMalloc(&e->start, (cells + 1) * 4));
  ENS_TRY(hipMalloc(&e->bsum, ((cells + kScanTile - 1) / kScanTile + 1) * 4));
  ENS_TRY(hipMemsetAsync(e->count, 0, (cells + 1) * 4, e->stream));
  // transpose host records (AoS of ncols floats) into columns, ids = index in member
  std::vector<char> colbuf(N * (size_t) e->elem);
  std::vector<int32_t> ids(N);
  for (int c = 0; c < e->ncols; c++) {
    for (size_t i = 0; i < total; i++) {
      std::memcpy(&colbuf[i * (size_t) e->elem], &host[(i * (size_t) e->ncols + (size_t) c) * (size_t) e->elem],
                  (size_t) e->elem);
    }
    ENS_TRY(hipMemcpy(e->col[0][c], colbuf.data(), total * (size_t) e->elem, hipMemcpyHostToDevice));
  }
  for (const Member &mb : e->mem) {
    for (int i = 0; i < mb.n; i++) ids[mb.off + (uint32_t) i] = i;
  }
  ENS_TRY(hipMemcpy(e->id[0], ids.data(), total * 4, hipMemcpyHostToDevice));
  e->cur = 0;
  e->id_cur = 0;
  // device tables
  std::vector<uint32_t> off(M + 1), cb(M + 1), tb(M + 1);
  std::vector<GridView> grids(M);
  for (int m = 0; m < M; m++) {
    off[m] = e->mem[m].off;
    cb[m] = e->mem[m].cell_base;
    tb[m] = e->mem[m].tile_base;
    grids[m] = e->mem[m].g;
  }
  off[M] = (uint32_t) total;
  cb[M] = (uint32_t) cells;
  tb[M] = tiles;
  ENS_TRY(hipMalloc(&e->d_off, (M + 1) * 4));
  ENS_TRY(hipMalloc(&e->d_cell_base, (M + 1) * 4));
  ENS_TRY(hipMalloc(&e->d_tile_base, (M + 1) * 4));
  ENS_TRY(hipMalloc(&e->d_grids, M * sizeof(GridView)));
  ENS_TRY(hipMemcpy(e->d_off, off.data(), (M + 1) * 4, hipMemcpyHostToDevice));
  ENS_TRY(hipMemcpy(e->d_cell_base, cb.data(), (M + 1) * 4, hipMemcpyHostToDevice));
  ENS_TRY(hipMemcpy(e->d_tile_base, tb.data(), (M + 1) * 4, hipMemcpyHostToDevice));
  ENS_TRY(hipMemcpy(e->d_grids, grids.data(), M * sizeof(GridView), hipMemcpyHostToDevice));
  // member kernel arguments: after the grid gather the columns sit in buffer
  // 1 - cur; the step reads them and writes buffer cur
  const size_t asz = e->hooks->args_size(e->precision);
  for (int b = 0; b < 2; b++) {
    std::vector<char> args((size_t) M * asz);
    for (int m = 0; m < M; m++) {
      ColSet in{}, outc{};
      for (int c = 0; c < e->ncols; c++) {
        in.c[c] = e->col[1 - b][c];
        outc.c[c] = e->col[b][c];
      }
      GridView g = e->mem[m].g;
      g.start = e->start + e->mem[m].cell_base;
      e->hooks->fill_args(e->precision, &args[(size_t) m * asz], e->mem[m].params.data(), in, outc,
                          g, e->mem[m].near_t, e->mem[m].rc, e->mem[m].n, e->occupancy, e->tile);
    }
    ENS_TRY(hipMalloc(&e->d_args[b], args.size()));
    ENS_TRY(hipMemcpy(e->d_args[b], args.data(), args.size(), hipMemcpyHostToDevice));
  }
  e->timestep = 0;
  if (e->lists && total > 0) {
    int rc = init_lists(e);
    if (rc) return rc;
  }
  ENS_TRY(hipStreamSynchronize(e->stream));
  e->ready = true;
  return ABL_OK;
}

// One timestep: rebuild every member's grid (one sort over all cells), then
// the program's tiled step over all members' tiles.  With lists: the gated
// rebuild, then the step through the lists from the current buffers into the
// other ones.
static int issue_step(abl_ensemble *e) {
  const int n = (int) e->n;
  const int M = (int) e->mem.size();
  hipStream_t st = e->stream;
  const int b = e->cur;
  if (e->lists) {
    if (n > 0) {
      issue_rebuild(e);
      NListArgs nl = ens_list_args(e);
      // the step reads buffer b: the arguments whose input is 1 - (1 - b)
      e->hooks->launch_multi_list(e->precision, e->d_args[1 - b], e->d_tile_base, M,
                                  e->total_tiles, e->reach, e->occupancy, e->tile, e->block, st,
                                  nl, e->nl_mgate, 2);
      e->cur ^= 1;
    }
    e->timestep += 1;
    ENS_TRY(hipGetLastError());
    return ABL_OK;
  }
  e->timestep += 1;
  if (n > 0) {
    if (e->elem == 8) {
      hipLaunchKernelGGL(k_ens_bin<double>, grid_blocks(n, 256), 256, 0, st,
                         (const double *) e->col[b][0], (const double *) e->col[b][1], e->d_off,
                         e->d_grids, e->d_cell_base, M, e->key, e->count, n);
    } else {
      hipLaunchKernelGGL(k_ens_bin<float>, grid_blocks(n, 256), 256, 0, st,
                         (const float *) e->col[b][0], (const float *) e->col[b][1], e->d_off,
                         e->d_grids, e->d_cell_base, M, e->key, e->count, n);
    }
  }
  const int nc = (int) e->ncells;
  const int ntiles = (nc + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL((k_scan_tiles<uint32_t, false>), ntiles, kScanBlock, 0, st, e->count, e->start,
                     e->bsum, nc);
  hipLaunchKernelGGL(k_scan_totals, 1, 1024, 0, st, e->bsum, ntiles, e->start, nc);
  hipLaunchKernelGGL(k_scan_add, grid_blocks(nc, 256), 256, 0, st, e->start, e->bsum, nc);
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, grid_blocks(n, 256), 256, 0, st, e->key, e->count, e->start,
                       e->id[e->id_cur], e->tmp_idx, e->tmp_id, n);
    GatherCols gc{};
    gc.ncols = e->ncols;
    for (int c = 0; c < e->ncols; c++) {
      gc.src[c] = e->col[b][c];
      gc.dst[c] = e->col[1 - b][c];
      gc.elem[c] = e->elem;
    }
    hipLaunchKernelGGL(k_rank_gather, grid_blocks(n, 256), 256, 0, st, e->tmp_idx, e->tmp_id,
                       e->key, e->start, gc, e->id[1 - e->id_cur], n);
    e->id_cur ^= 1;
    // the step reads buffer 1 - b and writes buffer b: columns end where they began
    e->hooks->launch_multi(e->precision, e->d_args[b], e->d_tile_base, M, e->total_tiles,
                           e->reach, e->occupancy, e->tile, e->block, st);
  }
  ENS_TRY(hipGetLastError());
  return ABL_OK;
}

extern "C" int abl_ens_simulate(abl_ensemble *e, int steps) {
  if (!e || steps < 0) return ens_fail(ABL_ERR_ARG, "bad argument");
  if (!e->ready) return ens_fail(ABL_ERR_STATE, "abl_ens_init first");
  ENS_TRY(hipEventRecord(e->ev0, e->stream));
  int t = 0;
  if (e->use_graph && steps > 0) {
    // the id buffer (without lists) or the column buffers and the gates'
    // parity (with lists) alternate per step: two timesteps, captured from
    // an even start, end where they start
    auto even = [&] { return e->lists ? (e->timestep & 1u) == 0 : (e->id_cur & 1) == 0; };
    if (!e->exec) {
      if (even() && steps >= 2) {
        const uint64_t t0 = e->timestep;
        ENS_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = issue_step(e);
        if (!rc) rc = issue_step(e);
        hipError_t ce = hipStreamEndCapture(e->stream, &e->graph);
        e->timestep = t0;  // captured, not run
        if (rc) return rc;
        ENS_TRY(ce);
        ENS_TRY(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
      }
    }
    if (e->exec && even()) {
      for (; t + 2 <= steps; t += 2) {
        ENS_TRY(hipGraphLaunch(e->exec, e->stream));
        e->timestep += 2;
      }
    }
  }
  for (; t < steps; t++) {
    int rc = issue_step(e);
    if (rc) return rc;
  }
  ENS_TRY(hipEventRecord(e->ev1, e->stream));
  return ABL_OK;
}

extern "C" int abl_ens_synchronize(abl_ensemble *e, double *ms) {
  if (!e) return ens_fail(ABL_ERR_ARG, "bad argument");
  ENS_TRY(hipStreamSynchronize(e->stream));
  float f = 0.f;
  if (hipEventElapsedTime(&f, e->ev0, e->ev1) == hipSuccess) e->last_ms = f;
  if (ms) *ms = e->last_ms;
  return ABL_OK;
}

extern "C" int abl_ens_member_count(abl_ensemble *e, int member, size_t *count) {
  if (!e || member < 0 || member >= (int) e->mem.size() || !count)
    return ens_fail(ABL_ERR_ARG, "bad argument");
  *count = (size_t) e->mem[(size_t) member].n;
  return ABL_OK;
}

// Member's agents in its own id order (as a run of it alone downloads them).
extern "C" int abl_ens_download(abl_ensemble *e, int member, void *aos, size_t capacity,
                                size_t *count) {
  if (!e || member < 0 || member >= (int) e->mem.size()) return ens_fail(ABL_ERR_ARG, "bad argument");
  if (!e->ready) return ens_fail(ABL_ERR_STATE, "abl_ens_init first");
  const Member &mb = e->mem[(size_t) member];
  if (count) *count = (size_t) mb.n;
  if (mb.n == 0) return ABL_OK;
  if (!aos || capacity < (size_t) mb.n) return ens_fail(ABL_ERR_ARG, "buffer too small");
  const size_t rec = (size_t) e->ncols * (size_t) e->elem;
  void *dst = nullptr, **dcols = nullptr;
  ENS_TRY(hipStreamSynchronize(e->stream));
  ENS_TRY(hipMalloc(&dst, (size_t) mb.n * rec));
  ENS_TRY(hipMalloc((void **) &dcols, kMaxCols * sizeof(void *)));
  void *hcols[kMaxCols] = {};
  for (int c = 0; c < e->ncols; c++) hcols[c] = e->col[e->cur][c];
  ENS_TRY(hipMemcpy(dcols, hcols, sizeof hcols, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_ens_unpermute, grid_blocks(mb.n, 256), 256, 0, e->stream, (char *) dst,
                     (const void *const *) dcols, e->elem, e->ncols, e->id[e->id_cur], mb.off, mb.n);
  ENS_TRY(hipGetLastError());
  ENS_TRY(hipMemcpyAsync(aos, dst, (size_t) mb.n * rec, hipMemcpyDeviceToHost, e->stream));
  ENS_TRY(hipStreamSynchronize(e->stream));
  ENS_TRY(hipFree(dst));
  ENS_TRY(hipFree(dcols));
  return ABL_OK;
}

extern "C" int abl_ens_agents(abl_ensemble *e, size_t *total) {
  if (!e || !total) return ens_fail(ABL_ERR_ARG, "bad argument");
  *total = e->n;
  return ABL_OK;
}

extern "C" int abl_ens_destroy(abl_ensemble *e) {
  if (!e) return ens_fail(ABL_ERR_ARG, "bad argument");
  (void) hipStreamSynchronize(e->stream);
  free_all(e);
  if (e->ev0) (void) hipEventDestroy(e->ev0);
  if (e->ev1) (void) hipEventDestroy(e->ev1);
  if (e->stream) (void) hipStreamDestroy(e->stream);
  delete e;
  return ABL_OK;
}
