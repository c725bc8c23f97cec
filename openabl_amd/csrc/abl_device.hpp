// abl_device.hpp -- device-side building blocks shared by the runtime's
// grid pipeline and the step programs (gfx950, wave64).
//
// Exactness rules the step programs rely on (all code here is compiled with
// -ffp-contract=off, so a*b+c is never fused):
//   * the near test of the reference C backend is
//       (double) sqrtf((float)(dx*dx + dy*dy [+ dz*dz])) <= r
//     (src/backend/CPrinter.cpp:168-170, asset/c/libabl.h:156-168).  sqrtf is
//     correctly rounded and monotone, so that test equals
//       (float)(dx*dx + dy*dy [+ dz*dz]) <= T,   T = max{f : sqrtf(f) <= r},
//     with T found on the host; the device never takes a square root for it.
//   * where a step needs the distance value itself, sqrt_rn_f32() returns the
//     correctly rounded single-precision root: the double root of a float is
//     correctly rounded, and rounding it again to float is exact because
//     53 >= 2*24 + 2 (double rounding is innocuous for sqrt).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace abl {

constexpr int kMaxCols = 32;
constexpr int kWave = 64;

// Correctly rounded sqrtf.
__device__ __forceinline__ float sqrt_rn_f32(float f) {
  return (float) __builtin_sqrt((double) f);
}

// Uniform grid of one agent type: cell edge = granularity / cell_scale,
// widened by 2^-20 so float rounding in the near test can never reach past
// the stencil; `reach` cells are searched on each side.
struct GridView {
  double lo[3];
  double inv_cell;   // 1 / cell edge
  int n[3];          // cells per axis (n[2] == 1 in 2-D)
  int reach;
  int dim;
  const uint32_t *start;  // ncells + 1 offsets into the cell-sorted order
};

__device__ __forceinline__ int cell_coord(double p, double lo, double inv, int n) {
  double q = floor((p - lo) * inv);
  // !(q >= 0) also maps NaN to cell 0.
  int c = !(q >= 0.0) ? 0 : (q > (double) (n - 1) ? n - 1 : (int) q);
  return c;
}

// Square of the reference's vector difference nx.pos - in.pos in abl_float
// arithmetic, rounded to float as the sqrtf argument is.
template <typename F>
__device__ __forceinline__ float dist_sq_f32(F ax, F ay, F bx, F by) {
  F dx = ax - bx;
  F dy = ay - by;
  F s = dx * dx + dy * dy;
  return (float) s;
}
template <typename F>
__device__ __forceinline__ float dist_sq_f32(F ax, F ay, F az, F bx, F by, F bz) {
  F dx = ax - bx;
  F dy = ay - by;
  F dz = az - bz;
  F s = dx * dx + dy * dy + dz * dz;
  return (float) s;
}

// Compensated (Kahan) float sum for the FP32-compensated precision: the
// running error c is fed back into the next addend.  Builds use
// -ffp-contract=off and no reassociation, so the compensation survives.
struct KahanF {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double) s - (double) c; }
};

// Column pointers of one agent type for a fused copy src[slot] -> dst[slot'].
struct GatherCols {
  const void *src[kMaxCols];
  void *dst[kMaxCols];
  int elem[kMaxCols];
  int ncols;
};

__device__ __forceinline__ void gather_cols(const GatherCols &g, uint32_t src, uint32_t dst) {
  for (int c = 0; c < g.ncols; c++) {
    switch (g.elem[c]) {
      case 1: ((uint8_t *) g.dst[c])[dst] = ((const uint8_t *) g.src[c])[src]; break;
      case 4: ((uint32_t *) g.dst[c])[dst] = ((const uint32_t *) g.src[c])[src]; break;
      default: ((uint64_t *) g.dst[c])[dst] = ((const uint64_t *) g.src[c])[src]; break;
    }
  }
}

// Neighbour-list state of one agent type (abl_tile.hpp, tiled_list_build /
// tiled_step_list).
constexpr uint16_t kListOverflow = 0xFFFF;
// lane order entry of a tile without lists: position p takes own agent p,
// which walks the stencil
constexpr uint32_t kOwnOrder = 0xFFFFFFFFu;

struct NListArgs {
  // halo indices in rows of eight entries (16 bytes): row r of agent i is
  // idx[(r * stride + i) * 8 ...], so a wave's lanes read one row coalesced
  uint16_t *idx;        // [cap / 8][stride][8]
  uint16_t *len;        // [n] entries, or kListOverflow
  // per tile, written by the build at the slot of the tile's own agent p:
  // the own index of the tile's p-th agent in order of list rows (<< 16) and
  // its list length, and that agent's first list row
  uint32_t *perm;       // [n]
  uint4 *prow0;         // [n]
  double *pos0;         // [dim][n] positions at the build
  uint32_t *gate;       // [2] rebuild flags by timestep parity
  uint32_t *stats;      // [0]: longest list at the last build; [2]: gated rebuilds run
  int n, cap, parity;
  int stride;           // agents per row block of idx (>= n)
  float rc_list;        // (radius + skin) in cell units, widened
  double r2_list;       // (radius + skin)^2, widened
  double half_skin2;    // (skin / 2)^2
  // rebuild only: copy of the freshly placed agents back to the current buffers
  GatherCols copy;
  const int32_t *id_src;
  int32_t *id_dst;
  const uint32_t *skey_src;
  uint32_t *skey_dst;
  const uint32_t *start_src;
  uint32_t *start_dst;
  int ncells;
  int gated;            // 1: build only when gate[parity] is set
};

// Fast float reciprocal (1 ulp) for the FP32-compensated precision.
__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }

// Inclusive-sum scan across one wave64.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t u = __shfl_up(v, d, kWave);
    if (lane >= d) v += u;
  }
  return v;
}

// A block-uniform value read from LDS, made wave-uniform for the compiler
// (scalar register): branches and loops on it are then scalar, not per-lane.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inclusive sum within each row of 16 lanes (DPP row shifts: no LDS
// traffic).  Lanes shifted in from outside the row contribute 0.
__device__ __forceinline__ uint32_t row16_incl_scan(uint32_t v) {
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xF, 0xF, true);  // row_shr:1
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xF, 0xF, true);  // row_shr:2
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xF, 0xF, true);  // row_shr:4
  v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xF, 0xF, true);  // row_shr:8
  return v;
}

// Block-wide exclusive scan (blockDim.x a multiple of 64, at most 1024).
// Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds_waves,
                                                    uint32_t *total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) lds_waves[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < nwaves ? lds_waves[lane] : 0u;
    uint32_t wi = wave_incl_scan(w);
    if (lane < nwaves) lds_waves[lane] = wi - w;  // exclusive wave offsets
    if (lane == nwaves - 1) lds_waves[kWave] = wi;
  }
  __syncthreads();
  uint32_t res = incl - v + lds_waves[wave];
  *total = lds_waves[kWave];
  __syncthreads();
  return res;
}

}  // namespace abl

namespace abl {

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X (each with
// its own L2).  Remap so that every XCD works on one contiguous range of the
// cell-sorted agents: neighbours of a block's agents then sit in the same
// XCD's L2 instead of being fetched by all eight.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned x = b & 7u, k = b >> 3;
  const unsigned q = nb >> 3, rem = nb & 7u;
  return x * q + (x < rem ? x : rem) + k;
}

// Visit the candidate agents of a stencil in cell-sorted order: for each row
// of cells (fixed y, z) the x-range of cells is one contiguous slot range.
template <typename Visit>
__device__ __forceinline__ void for_each_candidate(const GridView &g, int cx, int cy, int cz,
                                                   int reach, Visit &&visit) {
  const int x0 = cx - reach < 0 ? 0 : cx - reach;
  const int x1 = cx + reach > g.n[0] - 1 ? g.n[0] - 1 : cx + reach;
  const int y0 = cy - reach < 0 ? 0 : cy - reach;
  const int y1 = cy + reach > g.n[1] - 1 ? g.n[1] - 1 : cy + reach;
  int z0 = 0, z1 = 0;
  if (g.dim == 3) {
    z0 = cz - reach < 0 ? 0 : cz - reach;
    z1 = cz + reach > g.n[2] - 1 ? g.n[2] - 1 : cz + reach;
  }
  for (int z = z0; z <= z1; z++) {
    for (int y = y0; y <= y1; y++) {
      const uint32_t row = ((uint32_t) z * (uint32_t) g.n[1] + (uint32_t) y) * (uint32_t) g.n[0];
      const uint32_t a = g.start[row + x0];
      const uint32_t b = g.start[row + x1 + 1];
      for (uint32_t j = a; j < b; j++) visit(j);
    }
  }
}

// Conservative x-interval of cells, relative to the agent's own cell, that
// can hold agents within `rc` cell units of an agent at fractional in-cell
// position (ux, uy[, uz]) when looking at the row of cells dy (, dz) away.
// Returns false when the whole row is out of reach.  Positions outside
// [0, 1) (agents clamped into edge cells) disable narrowing.
__device__ __forceinline__ bool row_interval(float ux, float uy, float uz, int dy, int dz,
                                             float rc, int reach, int *xa, int *xb) {
  const bool inside = ux >= 0.f && ux < 1.f && uy >= 0.f && uy < 1.f && uz >= 0.f && uz < 1.f;
  if (!inside) {
    *xa = -reach;
    *xb = reach;
    return true;
  }
  const float gy = dy > 0 ? (float) dy - uy : (dy < 0 ? uy + (float) (-dy) - 1.f : 0.f);
  const float gz = dz > 0 ? (float) dz - uz : (dz < 0 ? uz + (float) (-dz) - 1.f : 0.f);
  const float gap2 = gy * gy + gz * gz;
  // radius squared, enlarged well past the float error of gap2 (values <= ~16,
  // error ~1e-6) so every agent the exact test can accept stays inside
  const float margin = 1e-4f;
  const float rr2 = rc * rc * (1.f + 1e-5f) + margin;
  if (gap2 > rr2) return false;
  const float w = __builtin_sqrtf(rr2 - gap2) + margin;
  int a = (int) floorf(ux - w);
  int b = (int) floorf(ux + w);
  *xa = a < -reach ? -reach : a;
  *xb = b > reach ? reach : b;
  return *xa <= *xb;
}

// The stencil rows of cells around the agent at (px, py, pz) narrowed to
// the cells that can hold agents within rc cell units (row_interval): row(a,
// b) for each non-empty range [a, b) of the cell-sorted order, rows in the
// order of for_each_candidate.  Only cells that cannot hold a neighbour are
// skipped, so a visit that applies the exact near test to every candidate
// accepts the same agents in the same order.
template <typename Row>
__device__ __forceinline__ void for_each_candidate_row(const GridView &g, double px, double py,
                                                       double pz, float rc, Row &&row) {
  const int cx = cell_coord(px, g.lo[0], g.inv_cell, g.n[0]);
  const int cy = cell_coord(py, g.lo[1], g.inv_cell, g.n[1]);
  const int cz = g.dim == 3 ? cell_coord(pz, g.lo[2], g.inv_cell, g.n[2]) : 0;
  const float ux = (float) ((px - g.lo[0]) * g.inv_cell - (double) cx);
  const float uy = (float) ((py - g.lo[1]) * g.inv_cell - (double) cy);
  const float uz = g.dim == 3 ? (float) ((pz - g.lo[2]) * g.inv_cell - (double) cz) : 0.f;
  const int R = g.reach;
  const int zr = g.dim == 3 ? R : 0;
  for (int dz = -zr; dz <= zr; dz++) {
    const int z = cz + dz;
    if (z < 0 || (g.dim == 3 && z >= g.n[2])) continue;
    for (int dy = -R; dy <= R; dy++) {
      const int y = cy + dy;
      if (y < 0 || y >= g.n[1]) continue;
      int xa, xb;
      if (!row_interval(ux, uy, uz, dy, dz, rc, R, &xa, &xb)) continue;
      const int x0 = max(cx + xa, 0);
      const int x1 = min(cx + xb, g.n[0] - 1);
      if (x0 > x1) continue;
      const uint32_t base = ((uint32_t) z * (uint32_t) g.n[1] + (uint32_t) y) * (uint32_t) g.n[0];
      const uint32_t a = g.start[base + x0];
      const uint32_t b = g.start[base + x1 + 1];
      if (a < b) row(a, b);
    }
  }
}

// A column read through a buffer resource: the column's base address is
// scalar and each lane supplies a 32-bit byte offset, so one offset serves
// every column of a record (the record's columns are at the same offset) and
// a read past `bytes` returns zero.  The resource's last word is the one the
// gfx9 family takes for untyped (raw) buffer access.
struct ColBuf {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ ColBuf col_buf(const void *p, uint32_t bytes) {
  return {__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int) bytes, 0x00020000)};
}
template <typename F>
__device__ __forceinline__ F col_load(const ColBuf &b, uint32_t off) {
  if constexpr (sizeof(F) == 8) {
    return __builtin_bit_cast(F, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int) off, 0, 0));
  } else {
    return __builtin_bit_cast(F, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int) off, 0, 0));
  }
}

// ---- quotients with a shared divisor --------------------------------------------
// a / b correctly rounded from y = RN(1 / b) with three more operations:
// q = RN(a y) is within an ulp of a / b, the residual a - b q is exact in one
// fma, and RN(q + (a - b q) y) is a / b correctly rounded (Markstein's
// theorem: P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Develop. 34(1), 1990; the final
// step of Newton-Raphson division in J.-M. Muller et al., Handbook of
// Floating-Point Arithmetic).  The theorem needs the quotient and the
// residual clear of underflow and overflow: div_ok() is the magnitude range
// that guarantees it (it also excludes zeros, whose sign the residual step
// can lose, and non-finite values); outside it the IEEE division is used.
// Quotients that share a divisor then cost one division in all.
template <typename F>
struct DivRange;
template <>
struct DivRange<double> {
  static constexpr double lo = 0x1p-400, hi = 0x1p400;
};
template <>
struct DivRange<float> {
  static constexpr float lo = 0x1p-50f, hi = 0x1p50f;
};
template <typename F>
__device__ __forceinline__ bool div_ok(F x) {
  const F ax = x < (F) 0 ? -x : x;
  return ax >= DivRange<F>::lo && ax <= DivRange<F>::hi;
}
__device__ __forceinline__ double fma_rn(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_rn(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// a / b given y = RN(1 / b), valid when div_ok(a) and div_ok(b)
template <typename F>
__device__ __forceinline__ F div_by_recip(F a, F b, F y) {
  const F q = a * y;
  const F r = fma_rn(-b, q, a);
  return fma_rn(r, y, q);
}

// Same candidates in ascending id order (the reference's visiting order):
// ids ascend inside every cell, so a merge over the stencil's cells yields
// the global id order.
template <typename Visit>
__device__ __forceinline__ void for_each_candidate_by_id(const GridView &g, int cx, int cy, int cz,
                                                         int reach, const int32_t *ids,
                                                         Visit &&visit) {
  const int x0 = cx - reach < 0 ? 0 : cx - reach;
  const int x1 = cx + reach > g.n[0] - 1 ? g.n[0] - 1 : cx + reach;
  const int y0 = cy - reach < 0 ? 0 : cy - reach;
  const int y1 = cy + reach > g.n[1] - 1 ? g.n[1] - 1 : cy + reach;
  int z0 = 0, z1 = 0;
  if (g.dim == 3) {
    z0 = cz - reach < 0 ? 0 : cz - reach;
    z1 = cz + reach > g.n[2] - 1 ? g.n[2] - 1 : cz + reach;
  }
  constexpr int kMaxRuns = 27;
  uint32_t cur[kMaxRuns], end[kMaxRuns];
  int32_t head[kMaxRuns];
  int nr = 0;
  for (int z = z0; z <= z1; z++) {
    for (int y = y0; y <= y1; y++) {
      const uint32_t row = ((uint32_t) z * (uint32_t) g.n[1] + (uint32_t) y) * (uint32_t) g.n[0];
      for (int x = x0; x <= x1 && nr < kMaxRuns; x++) {
        uint32_t a = g.start[row + x], b = g.start[row + x + 1];
        if (a < b) {
          cur[nr] = a;
          end[nr] = b;
          head[nr] = ids[a];
          nr++;
        }
      }
    }
  }
  while (nr > 0) {
    int best = 0;
    for (int r = 1; r < nr; r++) best = head[r] < head[best] ? r : best;
    const uint32_t j = cur[best];
    visit(j);
    cur[best] = j + 1;
    if (j + 1 < end[best]) {
      head[best] = ids[j + 1];
    } else {
      nr--;
      cur[best] = cur[nr];
      end[best] = end[nr];
      head[best] = head[nr];
    }
  }
}

template <typename F>
__device__ __forceinline__ int3 cell_of(const GridView &g, F x, F y, F z) {
  int3 c;
  c.x = cell_coord((double) x, g.lo[0], g.inv_cell, g.n[0]);
  c.y = cell_coord((double) y, g.lo[1], g.inv_cell, g.n[1]);
  c.z = g.dim == 3 ? cell_coord((double) z, g.lo[2], g.inv_cell, g.n[2]) : 0;
  return c;
}

}  // namespace abl
