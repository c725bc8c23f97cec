// abl_pipeline.hpp -- the grid pipeline's scan and sort kernels, shared by
// the single-simulation runtime (abl_runtime.hip) and the ensemble
// (abl_ensemble.hip).
//   scan     exclusive prefix of per-cell counts (wave64 shuffles, 3 passes)
//   scatter  slot = start[c] + --count[c]   (count returns to zero)
//   rank     stable order inside each cell by id, fused with the gather of
//            every column into the other buffer
#pragma once

#include "abl_device.hpp"

namespace abl {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// Per tile: exclusive scan of count into start, tile total into bsum.
// INV scans 1 - count (the survivors of a removeCurrent() flag array).
template <typename T, bool INV>
static __global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const T *__restrict__ count,
                                                          uint32_t *__restrict__ start,
                                                          uint32_t *__restrict__ bsum, int ncells) {
  __shared__ uint32_t lds[kWave + 1];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    uint32_t c = base + k < ncells ? (uint32_t) count[base + k] : (INV ? 1u : 0u);
    v[k] = INV ? 1u - c : c;
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = block_excl_scan(sum, lds, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < ncells) start[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the tile totals (one block, any number of tiles).
static __global__ __launch_bounds__(1024) void k_scan_totals(uint32_t *__restrict__ bsum, int ntiles,
                                                     uint32_t *__restrict__ start, int ncells) {
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
                           int ncells) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) start[c] += bsum[c / kScanTile];
}

static __global__ void k_scatter(const uint32_t *__restrict__ key, uint32_t *__restrict__ count,
                          const uint32_t *__restrict__ start, const int32_t *__restrict__ id,
                          uint32_t *__restrict__ tmp_idx, int32_t *__restrict__ tmp_id, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  const uint32_t slot = start[k] + atomicSub(&count[k], 1u) - 1u;
  tmp_idx[slot] = (uint32_t) i;
  tmp_id[slot] = id[i];
}

struct GatherCols {
  const void *src[kMaxCols];
  void *dst[kMaxCols];
  int elem[kMaxCols];
  int ncols;
};

// Stable placement inside each cell by id, fused with the column gather.
static __global__ void k_rank_gather(const uint32_t *__restrict__ tmp_idx, const int32_t *__restrict__ tmp_id,
                              const uint32_t *__restrict__ key, const uint32_t *__restrict__ start,
                              GatherCols g, int32_t *__restrict__ id_out, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t src = tmp_idx[s];
  const int32_t my = tmp_id[s];
  const uint32_t k = key[src];
  const uint32_t a = start[k], b = start[k + 1];
  uint32_t rank = 0;
  for (uint32_t t = a; t < b; t++) rank += tmp_id[t] < my ? 1u : 0u;
  const uint32_t dst = a + rank;
  id_out[dst] = my;
  for (int c = 0; c < g.ncols; c++) {
    switch (g.elem[c]) {
      case 1: ((uint8_t *) g.dst[c])[dst] = ((const uint8_t *) g.src[c])[src]; break;
      case 4: ((uint32_t *) g.dst[c])[dst] = ((const uint32_t *) g.src[c])[src]; break;
      default: ((uint64_t *) g.dst[c])[dst] = ((const uint64_t *) g.src[c])[src]; break;
    }
  }
}


}  // namespace abl
