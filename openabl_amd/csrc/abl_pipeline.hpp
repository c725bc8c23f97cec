// abl_pipeline.hpp -- the grid pipeline's scan and sort kernels, shared by
// the single-simulation runtime (abl_runtime.hip) and the ensemble
// (abl_ensemble.hip).
//   scan     exclusive prefix of per-cell counts (wave64 shuffles): three
//            passes (k_scan_tiles/totals/add) or two (k_tile_totals,
//            k_scan_apply)
//   scatter  slot = start[c] + --count[c]   (count returns to zero)
//   rank     stable order inside each cell by id, fused with the gather of
//            every column into the other buffer
//   delta    re-sort of an already sorted order after a step (k_bin_delta,
//            k_place): atomics only for the agents that changed cell
#pragma once

#include "abl_device.hpp"

namespace abl {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;
static_assert(kScanItems % 4 == 0, "a thread's run of cells is read as 16-byte words");

// A thread's kScanItems consecutive cells from `base` (zero past ncells):
// 16-byte accesses when the run is whole and aligned, else one word each.
__device__ __forceinline__ bool whole_run(const uint32_t *p, int base, int ncells) {
  return base + kScanItems <= ncells && (reinterpret_cast<uintptr_t>(p + base) & 15u) == 0;
}
__device__ __forceinline__ void load_run(const uint32_t *__restrict__ count, int base, int ncells,
                                         uint32_t *v) {
  if (whole_run(count, base, ncells)) {
    const uint4 *q = reinterpret_cast<const uint4 *>(count + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) {
      const uint4 x = q[j];
      v[4 * j] = x.x;
      v[4 * j + 1] = x.y;
      v[4 * j + 2] = x.z;
      v[4 * j + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++) v[k] = base + k < ncells ? count[base + k] : 0u;
  }
}
__device__ __forceinline__ void store_run(uint32_t *__restrict__ start, int base, int ncells,
                                          const uint32_t *w) {
  if (whole_run(start, base, ncells)) {
    uint4 *q = reinterpret_cast<uint4 *>(start + base);
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) q[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++)
      if (base + k < ncells) start[base + k] = w[k];
  }
}

// Per tile: exclusive scan of count into start, tile total into bsum.
// INV scans 1 - count (the survivors of a removeCurrent() flag array).
// With a gate the sort kernels here run only while *gate is set.
template <typename T, bool INV>
static __global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const T *__restrict__ count,
                                                          uint32_t *__restrict__ start,
                                                          uint32_t *__restrict__ bsum, int ncells,
                                                          const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t lds[kWave + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  if constexpr (sizeof(T) == 4 && !INV) {
    load_run(reinterpret_cast<const uint32_t *>(count), base, ncells, v);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      uint32_t c = base + k < ncells ? (uint32_t) count[base + k] : (INV ? 1u : 0u);
      v[k] = INV ? 1u - c : c;
    }
  }
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) sum += v[k];
  uint32_t total;
  uint32_t pre = block_excl_scan(sum, lds, &total);
  uint32_t w[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    w[k] = pre;
    pre += v[k];
  }
  store_run(start, base, ncells, w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the tile totals (one block, any number of tiles).
static __global__ __launch_bounds__(1024) void k_scan_totals(uint32_t *__restrict__ bsum, int ntiles,
                                                     uint32_t *__restrict__ start, int ncells,
                                                     const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t lds[kWave + 1];
  uint32_t carry = 0;
  for (int base = 0; base < ntiles; base += blockDim.x) {
    const int t = base + threadIdx.x;
    uint32_t v = t < ntiles ? bsum[t] : 0u;
    uint32_t total;
    uint32_t pre = block_excl_scan(v, lds, &total);
    if (t < ntiles) bsum[t] = pre + carry;
    carry += total;
  }
  if (threadIdx.x == 0) start[ncells] = carry;
}

static __global__ void k_scan_add(uint32_t *__restrict__ start, const uint32_t *__restrict__ bsum,
                           int ncells, const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) start[c] += bsum[c / kScanTile];
}

static __global__ void k_scatter(const uint32_t *__restrict__ key, uint32_t *__restrict__ count,
                          const uint32_t *__restrict__ start, const int32_t *__restrict__ id,
                          uint32_t *__restrict__ tmp_idx, int32_t *__restrict__ tmp_id, int n,
                          const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  const uint32_t slot = start[k] + atomicSub(&count[k], 1u) - 1u;
  tmp_idx[slot] = (uint32_t) i;
  tmp_id[slot] = id[i];
}

// Stable placement inside each cell by id, fused with the column gather.
// Optionally records the sorted keys (skey_out) and restores the per-cell
// counts the scatter consumed (count_out, written by each cell's rank-0
// agent; empty cells are already zero).
static __global__ void k_rank_gather(const uint32_t *__restrict__ tmp_idx, const int32_t *__restrict__ tmp_id,
                              const uint32_t *__restrict__ key, const uint32_t *__restrict__ start,
                              GatherCols g, int32_t *__restrict__ id_out, int n,
                              uint32_t *__restrict__ skey_out = nullptr,
                              uint32_t *__restrict__ count_out = nullptr,
                              const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t src = tmp_idx[s];
  const int32_t my = tmp_id[s];
  const uint32_t k = key[src];
  const uint32_t a = start[k], b = start[k + 1];
  uint32_t rank = 0;
  // four reads in flight: cells of dense clusters hold hundreds of agents
  uint32_t t = a;
  for (; t + 4 <= b; t += 4) {
    const int32_t i0 = tmp_id[t], i1 = tmp_id[t + 1], i2 = tmp_id[t + 2], i3 = tmp_id[t + 3];
    rank += (i0 < my ? 1u : 0u) + (i1 < my ? 1u : 0u) + (i2 < my ? 1u : 0u) + (i3 < my ? 1u : 0u);
  }
  for (; t < b; t++) rank += tmp_id[t] < my ? 1u : 0u;
  const uint32_t dst = a + rank;
  id_out[dst] = my;
  if (skey_out) skey_out[dst] = k;
  if (count_out && rank == 0) count_out[k] = b - a;
  gather_cols(g, src, dst);
}

// ---- two-kernel scan of per-cell counts --------------------------------------
// k_tile_totals: the total of every tile of kScanTile cells.  k_scan_apply:
// each tile's exclusive scan plus the sum of the totals of the tiles before
// it (each block adds those up itself), and start[ncells].

static __global__ __launch_bounds__(kScanBlock) void k_tile_totals(const uint32_t *__restrict__ count,
                                                                   uint32_t *__restrict__ bsum,
                                                                   int ncells) {
  __shared__ uint32_t lds[kWave + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  load_run(count, base, ncells, v);
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) sum += v[k];
  uint32_t total;
  (void) block_excl_scan(sum, lds, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Scan tile `tile` of `ntiles` (block-cooperative, kScanBlock threads; lds
// holds kWave + 1 words).
__device__ __forceinline__ void scan_apply_tile(int tile, int ntiles,
                                                const uint32_t *__restrict__ count,
                                                const uint32_t *__restrict__ bsum,
                                                uint32_t *__restrict__ start, int ncells,
                                                uint32_t *lds) {
  // the totals of the tiles before this one, read four at a time (bsum is
  // 16-byte aligned) with several reads in flight per thread: on a grid of
  // millions of cells a block adds up thousands of them
  uint32_t part = 0;
  const int quads = tile >> 2;
  const uint4 *b4 = reinterpret_cast<const uint4 *>(bsum);
#pragma unroll 4
  for (int t = threadIdx.x; t < quads; t += blockDim.x) {
    const uint4 q = b4[t];
    part += (q.x + q.y) + (q.z + q.w);
  }
  if ((int) threadIdx.x < (tile & 3)) part += bsum[(quads << 2) + threadIdx.x];
  uint32_t carry;
  (void) block_excl_scan(part, lds, &carry);
  const int base = tile * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  load_run(count, base, ncells, v);
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) sum += v[k];
  uint32_t total;
  uint32_t pre = block_excl_scan(sum, lds, &total) + carry;
  uint32_t w[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    w[k] = pre;
    pre += v[k];
  }
  store_run(start, base, ncells, w);
  if (tile == ntiles - 1 && threadIdx.x == 0) start[ncells] = carry + total;
}

static __global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t *__restrict__ count,
                                                                  const uint32_t *__restrict__ bsum,
                                                                  uint32_t *__restrict__ start,
                                                                  int ncells,
                                                                  const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t lds[kWave + 1];
  scan_apply_tile((int) blockIdx.x, (int) gridDim.x, count, bsum, start, ncells, lds);
}

// ---- delta re-sort -------------------------------------------------------------
// Between two builds on the same grid the stored order is sorted by (cell,
// id) and few agents change cell.  k_bin_delta computes every slot's new
// cell; an agent whose cell changed (a mover) moves one count (and tile
// total) from its old cell to its new one, tags both cells dirty for this
// build generation and pushes itself on its new cell's mover list.  k_place
// then puts every agent at its place in the (cell, id) order: in a clean cell
// the agents and their order are unchanged, so an agent keeps its offset from
// the cell's start; in a dirty cell its place is its rank by id among the
// cell's stayers and incoming movers.  The result is the full sort's.

struct MoverLists {
  uint32_t *dirty;                 // per cell: generation of the last build that touched it
  unsigned long long *head;        // per cell: (generation << 32) | first mover, newest first
  unsigned long long *next;        // per mover (numbered by its slot): the head it replaced
};

// table[idx] += v for every active lane with `on`, one atomic per distinct
// index among them (v is the same in every lane).  A wave holds consecutive
// agents of the cell order, so its movers leave from one or two cells and
// cross into a few scan tiles: per-lane atomics there all hit the same words.
__device__ __forceinline__ void wave_add(uint32_t *table, uint32_t idx, bool on, uint32_t v) {
  uint64_t left = __builtin_amdgcn_ballot_w64(on);
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  while (left != 0) {
    const int leader = __builtin_ctzll(left);
    const uint32_t li = (uint32_t) __builtin_amdgcn_readlane((int) idx, leader);
    const bool mine = on && idx == li;
    const uint64_t same = __builtin_amdgcn_ballot_w64(mine);
    if (lane == (uint32_t) leader) atomicAdd(&table[li], v * (uint32_t) __builtin_popcountll(same));
    left &= ~same;
  }
}

template <typename P>
__device__ __forceinline__ void bin_delta_item(int i, const P *__restrict__ px,
                                               const P *__restrict__ py, const P *__restrict__ pz,
                                               const GridView &g, const uint32_t *__restrict__ skey,
                                               uint32_t *__restrict__ key,
                                               uint32_t *__restrict__ count,
                                               uint32_t *__restrict__ bsum, const MoverLists &ml,
                                               uint32_t gen, int n) {
  const int3 c = cell_of<P>(g, px[i], py[i], g.dim == 3 ? pz[i] : (P) 0);
  const uint32_t k = ((uint32_t) c.z * (uint32_t) g.n[1] + (uint32_t) c.y) * (uint32_t) g.n[0] +
                     (uint32_t) c.x;
  key[i] = k;
  const uint32_t o = skey[i];
  if (k != o) {
    atomicAdd(&count[k], 1u);
    // leaving: few distinct cells per wave; tile totals change only for a
    // mover that crosses into another scan tile
    wave_add(count, o, true, ~0u);
    const bool cross = k / kScanTile != o / kScanTile;
    wave_add(bsum, k / kScanTile, cross, 1u);
    wave_add(bsum, o / kScanTile, cross, ~0u);
    ml.dirty[k] = gen;
    ml.dirty[o] = gen;
    // a mover is numbered by its slot (unique, below n, no shared counter;
    // the numbering is free, k_place ranks a cell's agents by id)
    const uint32_t m = (uint32_t) i;
    ml.next[m] = atomicExch(&ml.head[k], ((unsigned long long) gen << 32) | m);
  }
}

template <typename P>
static __global__ void k_bin_delta(const P *__restrict__ px, const P *__restrict__ py,
                                   const P *__restrict__ pz, GridView g,
                                   const uint32_t *__restrict__ skey, uint32_t *__restrict__ key,
                                   uint32_t *__restrict__ count, uint32_t *__restrict__ bsum,
                                   MoverLists ml, uint32_t gen, int n,
                                   const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;  // neighbour-list mode: no rebuild this timestep
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bin_delta_item<P>(i, px, py, pz, g, skey, key, count, bsum, ml, gen, n);
}

__device__ __forceinline__ void place_item(int s, const uint32_t *__restrict__ key,
                                           const uint32_t *__restrict__ skey_in,
                                           uint32_t *__restrict__ skey_out,
                                           const uint32_t *__restrict__ start_old,
                                           const uint32_t *__restrict__ start_new,
                                           const MoverLists &ml, uint32_t gen, const GatherCols &g,
                                           const int32_t *__restrict__ id_in,
                                           int32_t *__restrict__ id_out, int n) {
  if (s >= n) return;
  const uint32_t k = key[s];
  uint32_t dst;
  if (k == skey_in[s] && ml.dirty[k] != gen) {
    dst = start_new[k] + ((uint32_t) s - start_old[k]);
  } else {
    const int32_t my = id_in[s];
    // stayers with a smaller id: the cell's stored range [a, b) is in id
    // order, so they are the stayers before position p -- the agent's own
    // slot for a stayer, else the first position holding a larger id
    const uint32_t a = start_old[k], b = start_old[k + 1];
    uint32_t p = (uint32_t) s;
    if (k != skey_in[s]) {
      uint32_t lo = a, hi = b;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (id_in[mid] < my) lo = mid + 1;
        else hi = mid;
      }
      p = lo;
    }
    uint32_t rank = 0;
    uint32_t t = a;
    for (; t + 4 <= p; t += 4)  // four reads in flight
      rank += (key[t] == k ? 1u : 0u) + (key[t + 1] == k ? 1u : 0u) + (key[t + 2] == k ? 1u : 0u) +
              (key[t + 3] == k ? 1u : 0u);
    for (; t < p; t++) rank += key[t] == k ? 1u : 0u;
    // incoming movers of this generation (each pushed once: the list ends)
    unsigned long long h = ml.head[k];
    while ((uint32_t) (h >> 32) == gen) {
      const uint32_t m = (uint32_t) h;
      rank += id_in[m] < my ? 1u : 0u;  // mover m is the agent at slot m
      h = ml.next[m];
    }
    dst = start_new[k] + rank;
  }
  if (dst >= (uint32_t) n) return;  // unreachable while the counts are consistent
  skey_out[dst] = k;
  id_out[dst] = id_in[s];
  gather_cols(g, (uint32_t) s, dst);
}

// ---- the gated re-sort in one launch --------------------------------------------
// In neighbour-list mode the re-sort runs only on timesteps whose gate the
// previous step set.  Its kernels (bin, scan, place, then the list build)
// can run as one launch, so a timestep without a rebuild costs one launch that
// returns at once instead of four.
//
// The phases are ordered by a work queue, not by grid barriers.  The launch's
// work is a sequence of items -- bin chunks, scan tiles, place chunks, then
// list-build tiles -- that blocks claim one at a time with an atomic counter,
// in increasing order.  Before it works on an item of phase p a block waits
// until every item of phase p - 1 is finished.  Those items were all claimed
// before this one, so each is held by a block that is already running and
// stays resident until it is done; phase-0 items wait for nothing.  By
// induction over the phases every wait ends whether or not the launch's
// blocks fit on the device at once, and whatever else runs beside it (other
// simulations' launches on other streams).  A block that finds the queue
// empty leaves; the last one out resets the queue for the next launch.

// Work-queue state of one agent type (device memory, zero between launches).
struct RebuildQueue {
  uint32_t next;     // items claimed
  uint32_t done[4];  // items finished, per phase
  uint32_t exited;   // blocks that found the queue empty
  uint32_t error;    // a wait ran past its spin limit (a safety net: by the
                     // argument above no wait can, unless the limit is forced)
  uint32_t pad;
};
static_assert(sizeof(RebuildQueue) == 32, "the queue occupies eight words of the gate block");

// Everything the re-sort phases read and write.
struct DeltaArgs {
  const void *px, *py, *pz;   // positions (current buffers)
  int pos_elem;               // 8 (double) or 4 (float)
  GridView g;                 // grid the agents are binned on
  const uint32_t *skey;       // sorted keys of the stored order
  uint32_t *key, *count, *bsum;
  MoverLists ml;
  uint32_t gen;
  int n;
  const uint32_t *start_old;
  uint32_t *start_new;
  int ncells, nscan;          // cells, scan tiles
  uint32_t *skey_out;
  GatherCols gc;              // current -> other buffers
  const int32_t *id_in;
  int32_t *id_out;
  RebuildQueue *q;
  uint32_t spins;             // spin limit of one phase wait (kPhaseSpins)
  unsigned blocks;            // launch size (0: as many blocks as the CUs hold at once)
  int *launched;              // host: set to 1 when the launch function launched it
};

// Spin limit of one phase wait (each spin sleeps ~0.1 us plus a load): far
// above any phase of a rebuild.
constexpr uint32_t kPhaseSpins = 1u << 22;

// Items per phase of a launch with blocks of 256 threads.
struct QueuePlan {
  uint32_t end[4];  // exclusive end of each phase's items
  __device__ __host__ int phase_of(uint32_t item) const {
    return item < end[0] ? 0 : item < end[1] ? 1 : item < end[2] ? 2 : 3;
  }
  __device__ __host__ uint32_t first(int p) const { return p == 0 ? 0u : end[p - 1]; }
  __device__ __host__ uint32_t items(int p) const { return end[p] - first(p); }
};

// Claim the block's next item (block-uniform result).
__device__ __forceinline__ uint32_t queue_claim(RebuildQueue *q) {
  __shared__ uint32_t claimed;
  __syncthreads();  // every thread has read the previous claim
  if (threadIdx.x == 0)
    claimed = __hip_atomic_fetch_add(&q->next, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return claimed;
}

// Wait until `count` items of phase p are finished.  Thread 0 waits with an
// acquire load; after the block barrier every thread acquires at agent scope
// too, so its reads see the finished phase's writes from every XCD.  Returns
// false when the wait ran past the spin limit or another block's did.
__device__ __forceinline__ bool queue_wait(RebuildQueue *q, int p, uint32_t count, uint32_t spins) {
  __shared__ int ok;
  if (threadIdx.x == 0) {
    int good = 1;
    for (uint32_t spin = 0;; spin++) {
      if (__hip_atomic_load(&q->done[p], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= count) break;
      if (spin >= spins ||
          __hip_atomic_load(&q->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        __hip_atomic_fetch_or(&q->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    ok = good;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok != 0;
}

// Mark the block's item of phase p finished: every thread releases its writes
// at agent scope before the block barrier, then thread 0 counts the item.
__device__ __forceinline__ void queue_finish(RebuildQueue *q, int p) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add(&q->done[p], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The last block out of a launch zeroes the queue for the next one.  Every
// block leaves only after finishing the items it claimed, so when the last
// one leaves every item is done.
__device__ __forceinline__ void queue_exit(RebuildQueue *q) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t e =
        __hip_atomic_fetch_add(&q->exited, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (e == gridDim.x - 1u) {
      __hip_atomic_store(&q->next, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int p = 0; p < 4; p++)
        __hip_atomic_store(&q->done[p], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&q->exited, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// One item of the re-sort's phases 0-2 (blockDim.x == kScanBlock == 256):
// bin chunk, scan tile, or place chunk `k` of its phase.
template <typename P>
__device__ __forceinline__ void delta_item(const DeltaArgs &d, int phase, uint32_t k,
                                           uint32_t *lds) {
  if (phase == 0) {
    const int i = (int) (k * blockDim.x + threadIdx.x);
    if (i < d.n)
      bin_delta_item<P>(i, (const P *) d.px, (const P *) d.py, (const P *) d.pz, d.g, d.skey,
                        d.key, d.count, d.bsum, d.ml, d.gen, d.n);
  } else if (phase == 1) {
    scan_apply_tile((int) k, d.nscan, d.count, d.bsum, d.start_new, d.ncells, lds);
  } else {
    place_item((int) (k * blockDim.x + threadIdx.x), d.key, d.skey, d.skey_out, d.start_old,
               d.start_new, d.ml, d.gen, d.gc, d.id_in, d.id_out, d.n);
  }
}

static __global__ void k_place(const uint32_t *__restrict__ key, const uint32_t *__restrict__ skey_in,
                               uint32_t *__restrict__ skey_out,
                               const uint32_t *__restrict__ start_old,
                               const uint32_t *__restrict__ start_new, MoverLists ml, uint32_t gen,
                               GatherCols g, const int32_t *__restrict__ id_in,
                               int32_t *__restrict__ id_out, int n,
                               const uint32_t *__restrict__ gate = nullptr) {
  if (gate && *gate == 0u) return;
  place_item((int) (blockIdx.x * blockDim.x + threadIdx.x), key, skey_in, skey_out, start_old,
             start_new, ml, gen, g, id_in, id_out, n);
}


}  // namespace abl
