// abl_tile.hpp -- LDS-tiled neighbour iteration for the step programs.
//
// One block owns a tile of TX x TY (x TZ) grid cells.  It stages, for the
// tile plus a halo of `reach` cells on every side, the neighbour type's
// positions and the other members the step reads, into LDS as one record
// (array of structs) per agent: every halo row of cells is one contiguous
// range of the cell-sorted columns, so the staging loads are coalesced.  Each
// thread then takes one of the tile's own agents and walks its stencil out of
// LDS.  A halo that does not fit the block's LDS budget (very dense cells)
// falls back to reading the same candidates from global memory, so results
// never depend on the budget.
//
// Per-row narrowing: for each row of cells of the stencil the thread computes,
// in cell units and conservatively (a 1e-4-cell margin on float arithmetic),
// the x-interval that can hold agents within the radius, and visits only those
// cells.  With cells of granularity / 2 this halves the candidates of a 3x3
// stencil of granularity-sized cells.  The exact near test is applied to
// every candidate, so narrowing never changes which agents are accepted.
#pragma once

#include <atomic>
#include <type_traits>

#include "abl_device.hpp"
#include "abl_pipeline.hpp"
#include "abl_runtime.hpp"

namespace abl {

template <int DIM, int TX, int TY, int TZ, int RM>
struct TileTables {
  static constexpr int HX = TX + 2 * RM;
  static constexpr int HY = TY + 2 * RM;
  static constexpr int HZ = DIM == 3 ? TZ + 2 * RM : 1;
  static constexpr int MAXROWS = HY * HZ;
  static constexpr int OWNROWS = TY * (DIM == 3 ? TZ : 1);
  int h0[3], hn[3], t0[3], tn[3];
  int nrows, total, own_total, overflow;
  uint32_t hstart[MAXROWS];            // global slot of each halo row's first agent
  uint32_t lbase[MAXROWS + 1];         // LDS record index of each halo row's first agent
  uint32_t ownbase[OWNROWS + 1];       // prefix of the tile's own agents per own row
  uint32_t ownrow[OWNROWS];            // halo row index of each own row
  uint32_t lcell[MAXROWS * (HX + 1)];  // global slot of each halo cell's start

  // LDS record index of halo cell hx of halo row `row` (hx = hn[0]: one past)
  __device__ __forceinline__ uint32_t cell_lds(int row, int hx) const {
    return lbase[row] + (lcell[row * (HX + 1) + hx] - hstart[row]);
  }
  // global slot of LDS record l in halo row r
  __device__ __forceinline__ uint32_t global_of(int r, uint32_t l) const {
    return hstart[r] + (l - lbase[r]);
  }

  // Geometry of tile `tile` (row-major over the tile grid): its cells
  // [T0, T0 + TN) and its halo [H0, H0 + HN) on each axis.
  struct Geom {
    int T0[3], TN[3], H0[3], HN[3];
  };
  __device__ static Geom geom(const GridView &g, int reach, unsigned tile) {
    Geom G;
    const int ntx = (g.n[0] + TX - 1) / TX;
    const int nty = (g.n[1] + TY - 1) / TY;
    int b[3];
    b[0] = (int) (tile % (unsigned) ntx);
    b[1] = (int) ((tile / (unsigned) ntx) % (unsigned) nty);
    b[2] = DIM == 3 ? (int) (tile / ((unsigned) ntx * (unsigned) nty)) : 0;
    const int T[3] = {TX, TY, DIM == 3 ? TZ : 1};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (d < DIM) {
        G.T0[d] = b[d] * T[d];
        G.TN[d] = min(T[d], g.n[d] - G.T0[d]);
        G.H0[d] = max(G.T0[d] - reach, 0);
        G.HN[d] = min(G.T0[d] + G.TN[d] + reach, g.n[d]) - G.H0[d];
      } else {
        G.T0[d] = G.H0[d] = 0;
        G.TN[d] = G.HN[d] = 1;
      }
    }
    return G;
  }

  // Global offset (cell-sorted slot) of cell hx (0..HN[0]) of halo row r.
  __device__ static uint32_t halo_start(const GridView &g, const Geom &G, int r, int hx) {
    int hy = r, hz = 0;
    if constexpr (DIM == 3) {
      hy = r % G.HN[1];
      hz = r / G.HN[1];
    }
    const uint32_t grow = ((uint32_t) (G.H0[2] + hz) * (uint32_t) g.n[1] + (uint32_t) (G.H0[1] + hy)) *
                          (uint32_t) g.n[0];
    return g.start[grow + G.H0[0] + hx];
  }

  // Block-cooperative set-up for tile `tile` (row-major over the tile grid).
  // `cap` is the LDS record capacity; overflow is set when the halo exceeds it.
  // One round of global reads: every thread derives the tile geometry itself,
  // the cell offsets of all halo cells are read at once, and the first wave
  // forms the per-row prefixes from them (blockDim.x >= 64).  Two barriers.
  __device__ void setup(const GridView &g, int reach, unsigned tile, int cap) {
    const Geom G = geom(g, reach, tile);
    const int *T0 = G.T0, *TN = G.TN, *H0 = G.H0, *HN = G.HN;
    const int nr = HN[1] * HN[2];
    const int w = HN[0] + 1;
    if (threadIdx.x == 0) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        t0[d] = T0[d];
        tn[d] = TN[d];
        h0[d] = H0[d];
        hn[d] = HN[d];
      }
      nrows = nr;
    }
    // global offsets of every halo cell (cell_lds converts them on read);
    // entry e of the table is cell e % (HX + 1) of halo row e / (HX + 1)
    for (int e = threadIdx.x; e < nr * (HX + 1); e += blockDim.x) {
      const int r = e / (HX + 1), hx = e % (HX + 1);
      if (hx < w) lcell[e] = halo_start(g, G, r, hx);
    }
    __syncthreads();
    const int oy = T0[1] - H0[1], oz = T0[2] - H0[2], ox = T0[0] - H0[0];
    const int nq = TN[1] * TN[2];
    if (nr <= 16 && nq <= 16) {
      // one wave: lanes 0-15 the halo rows (first slot, agent count), lanes
      // 16-31 the tile's own rows, each group one 16-lane prefix
      if (threadIdx.x < (unsigned) kWave) {
        const int lane = (int) threadIdx.x;
        uint32_t cnt = 0;
        if (lane < nr) {
          const uint32_t a = lcell[lane * (HX + 1)];
          hstart[lane] = a;
          cnt = lcell[lane * (HX + 1) + HN[0]] - a;
        } else if (lane >= 16 && lane < 16 + nq) {
          const int q = lane - 16;
          const int r = DIM == 3 ? (oz + q / TN[1]) * HN[1] + (oy + q % TN[1]) : oy + q;
          ownrow[q] = (uint32_t) r;
          cnt = lcell[r * (HX + 1) + ox + TN[0]] - lcell[r * (HX + 1) + ox];
        }
        const uint32_t incl = row16_incl_scan(cnt);
        if (lane < nr) lbase[lane + 1] = incl;
        if (lane >= 16 && lane < 16 + nq) ownbase[lane - 16 + 1] = incl;
        const uint32_t tot = (uint32_t) __builtin_amdgcn_readlane((int) incl, 15);
        const uint32_t otot = (uint32_t) __builtin_amdgcn_readlane((int) incl, 31);
        if (lane == 0) {
          lbase[0] = 0;
          total = (int) tot;
          // one record more than the halo: the far-away pad record of the lists
          overflow = (int) tot >= cap ? 1 : 0;
          ownbase[0] = 0;
          own_total = (int) otot;
        }
      }
    } else if (threadIdx.x < (unsigned) kWave) {
      const int lane = (int) threadIdx.x;
      // halo rows: first slot, agent count, prefix
      uint32_t carry = 0;
      for (int r0 = 0; r0 < nr; r0 += kWave) {
        const int r = r0 + lane;
        uint32_t cnt = 0;
        if (r < nr) {
          const uint32_t a = lcell[r * (HX + 1)];
          hstart[r] = a;
          cnt = lcell[r * (HX + 1) + HN[0]] - a;
        }
        const uint32_t incl = wave_incl_scan(cnt);
        if (r < nr) lbase[r + 1] = carry + incl;
        carry += (uint32_t) __shfl((int) incl, kWave - 1, kWave);
      }
      if (lane == 0) {
        lbase[0] = 0;
        total = (int) carry;
        // one record more than the halo: the far-away pad record of the lists
        overflow = (int) carry >= cap ? 1 : 0;
      }
      // the tile's own rows: their halo row and a prefix of their agents
      uint32_t ocarry = 0;
      for (int q0 = 0; q0 < nq; q0 += kWave) {
        const int q = q0 + lane;
        uint32_t cnt = 0;
        if (q < nq) {
          const int r = DIM == 3 ? (oz + q / TN[1]) * HN[1] + (oy + q % TN[1]) : oy + q;
          ownrow[q] = (uint32_t) r;
          cnt = lcell[r * (HX + 1) + ox + TN[0]] - lcell[r * (HX + 1) + ox];
        }
        const uint32_t incl = wave_incl_scan(cnt);
        if (q < nq) ownbase[q + 1] = ocarry + incl;
        ocarry += (uint32_t) __shfl((int) incl, kWave - 1, kWave);
      }
      if (lane == 0) {
        ownbase[0] = 0;
        own_total = (int) ocarry;
      }
    }
    __syncthreads();
  }

  // Halo row of LDS record k (binary search over lbase).
  __device__ __forceinline__ int row_of_record(uint32_t k) const {
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (lbase[mid] <= k) lo = mid; else hi = mid - 1;
    }
    return lo;
  }

  // Own agent k of the tile: its halo row and LDS record index.
  __device__ __forceinline__ void own(int k, int *row, uint32_t *rec) const {
    const int nq = tn[1] * tn[2];
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ownbase[mid] <= (uint32_t) k) lo = mid; else hi = mid - 1;
    }
    const int r = (int) ownrow[lo];
    *row = r;
    *rec = cell_lds(r, t0[0] - h0[0]) + ((uint32_t) k - ownbase[lo]);
  }
};

// Tiles of TX x TY (x TZ) cells covering the grid.
template <int TX, int TY, int TZ, int DIM>
__host__ __device__ inline unsigned tile_count(const GridView &g) {
  unsigned n = (unsigned) (((g.n[0] + TX - 1) / TX) * ((g.n[1] + TY - 1) / TY));
  if (DIM == 3) n *= (unsigned) ((g.n[2] + TZ - 1) / TZ);
  return n;
}

}  // namespace abl

namespace abl {

// Generic LDS-tiled step kernel for a step whose near loop runs over its own
// agent type.  The policy M supplies:
//   using F; static constexpr int DIM, NS;   NS = abl_float values per record
//   struct Args  (with GridView g; float rc; int cap;)
//   static void load(const Args &, uint32_t slot, F *rec)     -- NS values
//   struct Acc   { void visit(const Args &, const F *own, const F *nb); }
//   static void begin(const Args &, Acc &, const F *own)      -- before the loop
//   static void finish(const Args &, Acc &, const F *own, uint32_t slot, double *np)
//       -- writes the agent's outputs to slot; np (when not null) receives
//          the new position in FP64
// Records are the first DIM values = position, then the members the near
// loop reads.  Candidates are visited row by row of the (narrowed) stencil,
// each row in cell-sorted order: the same order as for_each_candidate.
//
// A policy that declares `static constexpr bool kRelative = true` keeps FP64
// positions in HBM but stages them in LDS as float offsets from the tile's
// centre (FP32-compensated precision).  It then supplies instead
//   static void load(const Args &, uint32_t slot, F *rec, const double org[3])
//   static void position(const Args &, uint32_t slot, double p[3])  -- FP64 position
//   static void begin(const Args &, Acc &, const F *own, const double p[3], const double org[3])
//   static void finish(const Args &, Acc &, const F *own, uint32_t slot, const double p[3],
//                      double *np)
// begin() receives the agent's FP64 position and the origin, so the policy
// can remove the rounding of the agent's own float record from every offset
// (otherwise one systematic error shared by all of its neighbour terms).
// Waves per SIMD the list step kernel is compiled for (its register budget):
// a policy may declare `static constexpr int kListWaves`; otherwise the
// compiler's own choice stands.
template <class M, class = void>
struct policy_list_waves {
  static constexpr int value = 1;
};
template <class M>
struct policy_list_waves<M, decltype((void) M::kListWaves)> {
  static constexpr int value = M::kListWaves;
};

template <class M, class = void>
struct policy_relative {
  static constexpr bool value = false;
};
template <class M>
struct policy_relative<M, decltype((void) M::kRelative)> {
  static constexpr bool value = M::kRelative;
};

// A policy whose agent's whole step runs as one function (the generated
// programs): M::agent(a, gi, walk, np) runs agent gi's step, handing the
// step's near loop `walk`, which calls visit(nb) for every candidate record
// in order (visit returns false to stop), and leaves the new position in np.
template <class M, class = void>
struct policy_agent {
  static constexpr bool value = false;
};
template <class M>
struct policy_agent<M, decltype((void) M::kAgentStep)> {
  static constexpr bool value = M::kAgentStep;
};

template <class M, int TX, int TY, int TZ, int RM>
__device__ __forceinline__ void tiled_body(const typename M::Args &a, unsigned tile,
                                           unsigned char *smem);

template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256) void tiled_step(typename M::Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  tiled_body<M, TX, TY, TZ, RM>(a, xcd_block(blockIdx.x, gridDim.x), smem);
}

// Ensemble form: the tiles of many independent members in one launch.
// Member m owns tiles [tile_base[m], tile_base[m+1]) and has its own
// arguments (grid, parameters); the block's member is uniform, so its
// arguments are read through the scalar cache.
template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256) void tiled_step_multi(const typename M::Args *__restrict__ args,
                                                        const uint32_t *__restrict__ tile_base,
                                                        int members) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  int lo = 0, hi = members - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_base[mid] <= t) lo = mid; else hi = mid - 1;
  }
  tiled_body<M, TX, TY, TZ, RM>(args[lo], t - tile_base[lo], smem);
}

// A coordinate whose square overflows to +inf in F.
template <typename F>
constexpr double kFarAway = sizeof(F) == 8 ? 1e300 : 1e30;

// One tile's block state: tables, LDS records, overflow, relative origin.
template <class M, int TX, int TY, int TZ, int RM>
struct TileCtx {
  using F = typename M::F;
  using TT = TileTables<M::DIM, TX, TY, TZ, RM>;
  TT *T;
  F *rec;
  bool ovf;
  double org[3];
};

// Set up tile `tile` and stage its halo records (block-cooperative).
// issue(c) runs after the set-up and before the record loads, so global
// reads it starts are in flight together with them.
template <class M, int TX, int TY, int TZ, int RM, class Issue>
__device__ __forceinline__ TileCtx<M, TX, TY, TZ, RM> tile_stage(const typename M::Args &a,
                                                                 unsigned tile,
                                                                 unsigned char *smem,
                                                                 Issue &&issue) {
  using Ctx = TileCtx<M, TX, TY, TZ, RM>;
  using F = typename M::F;
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  Ctx c;
  c.T = reinterpret_cast<typename Ctx::TT *>(smem);
  c.rec = reinterpret_cast<F *>(smem + ((sizeof(typename Ctx::TT) + 15) & ~(size_t) 15));
  const GridView &g = a.g;
  c.T->setup(g, g.reach, tile, a.cap);
  c.ovf = uniform(c.T->overflow) != 0;
  c.org[0] = c.org[1] = c.org[2] = 0.0;
  if constexpr (policy_relative<M>::value) {
#pragma unroll
    for (int d = 0; d < DIM; d++)
      c.org[d] = g.lo[d] + ((double) c.T->t0[d] + 0.5 * (double) c.T->tn[d]) / g.inv_cell;
  }
  issue(c);
  if (!c.ovf) {
    // kBatch records per thread in flight: all their global reads are issued
    // before the first is written to LDS
    constexpr int kBatch = 4;
    const int total = uniform(c.T->total);
    constexpr int kGroup = 16;
    const int ngroups = (int) blockDim.x / kGroup;
    if (uniform(c.T->nrows) <= ngroups) {
      // a group of 16 lanes per halo row: the row is one contiguous range of
      // slots and of records, so no search for a record's row
      const int r = (int) threadIdx.x / kGroup;
      if (r < c.T->nrows) {
        const uint32_t lb = c.T->lbase[r], le = c.T->lbase[r + 1], hs = c.T->hstart[r];
        for (uint32_t k0 = lb + threadIdx.x % kGroup; k0 < le; k0 += kBatch * kGroup) {
          F rr[kBatch][NS];
#pragma unroll
          for (int u = 0; u < kBatch; u++) {
            const uint32_t k = k0 + (uint32_t) (u * kGroup);
            if (k < le) {
              if constexpr (policy_relative<M>::value) M::load(a, hs + (k - lb), rr[u], c.org);
              else M::load(a, hs + (k - lb), rr[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < kBatch; u++) {
            const uint32_t k = k0 + (uint32_t) (u * kGroup);
            if (k < le) {
#pragma unroll
              for (int v = 0; v < NS; v++) c.rec[(size_t) NS * k + v] = rr[u][v];
            }
          }
        }
      }
    } else {
      // more rows than groups: threads stride over the records
      for (int k0 = threadIdx.x; k0 < total; k0 += kBatch * (int) blockDim.x) {
        F r[kBatch][NS];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          const int k = k0 + u * (int) blockDim.x;
          if (k < total) {
            const uint32_t j = c.T->global_of(c.T->row_of_record((uint32_t) k), (uint32_t) k);
            if constexpr (policy_relative<M>::value) M::load(a, j, r[u], c.org);
            else M::load(a, j, r[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
          const int k = k0 + u * (int) blockDim.x;
          if (k < total) {
#pragma unroll
            for (int v = 0; v < NS; v++) c.rec[(size_t) NS * k + v] = r[u][v];
          }
        }
      }
    }
    // record `total`: a position no near test accepts (its squares overflow),
    // the entry neighbour lists are padded with
    if (threadIdx.x == 0) {
#pragma unroll
      for (int v = 0; v < NS; v++)
        c.rec[(size_t) NS * c.T->total + v] = v < DIM ? (F) kFarAway<F> : (F) 0;
    }
  }
  __syncthreads();
  return c;
}

// Set-up and staging of a 2-D tile whose halo has at most 16 rows, for a
// block of 256 threads (16 groups of 16 lanes), with one barrier.  Every
// group reads the slot ranges of the halo rows and of the tile's own rows
// itself (lane j: row j) and forms their prefixes with 16-lane DPP scans in
// registers, so group r starts reading halo row r's records as soon as the
// cell offsets arrive, and own agent threadIdx.x is known at once:
// own(k, gi, l) runs for it (if it exists) with its slot and LDS record
// index, before the record reads, so reads it issues are in flight with them.  The LDS tables are
// written meanwhile.
template <class M, int TX, int TY, int TZ, int RM, class Own>
__device__ __forceinline__ TileCtx<M, TX, TY, TZ, RM> tile_stage_rows(const typename M::Args &a,
                                                                      unsigned tile,
                                                                      unsigned char *smem,
                                                                      Own &&own) {
  using Ctx = TileCtx<M, TX, TY, TZ, RM>;
  using TT = typename Ctx::TT;
  using F = typename M::F;
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  static_assert(DIM == 2 && TT::MAXROWS <= 16, "row-group staging: 2-D halos of <= 16 rows");
  constexpr int kBatch = 4;
  Ctx c;
  c.T = reinterpret_cast<TT *>(smem);
  c.rec = reinterpret_cast<F *>(smem + ((sizeof(TT) + 15) & ~(size_t) 15));
  c.org[0] = c.org[1] = c.org[2] = 0.0;
  const GridView &g = a.g;
  const typename TT::Geom G = TT::geom(g, g.reach, tile);
  const int nr = G.HN[1];
  const int w = G.HN[0] + 1;
  const int nq = G.TN[1];
  const int oy = G.T0[1] - G.H0[1], ox = G.T0[0] - G.H0[0];
  const int lane16 = (int) (threadIdx.x & 15u), grp = (int) (threadIdx.x >> 4);
  const uint32_t nx = (uint32_t) g.n[0];
  // one round of global reads: row ranges (per group) and halo cell offsets
  uint32_t a0 = 0, a1 = 0, o0 = 0, o1 = 0;
  if (lane16 < nr) {
    const uint32_t grow = (uint32_t) (G.H0[1] + lane16) * nx;
    a0 = g.start[grow + G.H0[0]];
    a1 = g.start[grow + G.H0[0] + G.HN[0]];
  }
  if (lane16 < nq) {
    const uint32_t grow = (uint32_t) (G.T0[1] + lane16) * nx;
    o0 = g.start[grow + G.T0[0]];
    o1 = g.start[grow + G.T0[0] + G.TN[0]];
  }
  // lcell entry e: cell e % (HX + 1) of halo row e / (HX + 1) (a constant
  // divisor: no integer division at run time)
  constexpr int LW = TT::HX + 1;
  const int ne = nr * LW;
  const int er = (int) threadIdx.x / LW, ex = (int) threadIdx.x % LW;
  const bool eok = (int) threadIdx.x < ne && ex < w;
  const uint32_t raw = eok ? TT::halo_start(g, G, er, ex) : 0u;
  const uint32_t cnt = a1 - a0, incl = row16_incl_scan(cnt);
  const uint32_t ocnt = o1 - o0, oincl = row16_incl_scan(ocnt);
  const uint32_t total = (uint32_t) __builtin_amdgcn_readlane((int) incl, 15);
  const uint32_t own_total = (uint32_t) __builtin_amdgcn_readlane((int) oincl, 15);
  c.ovf = (int) total >= a.cap;
  if constexpr (policy_relative<M>::value) {
#pragma unroll
    for (int d = 0; d < DIM; d++)
      c.org[d] = g.lo[d] + ((double) G.T0[d] + 0.5 * (double) G.TN[d]) / g.inv_cell;
  }
  // own agent threadIdx.x: its own row q (the first with oincl[q] > k) and
  // slot (the shuffles run on every lane: their sources must be active)
  {
    const uint32_t k = threadIdx.x;
    int q = 0;
    // (every 16-lane group holds the same prefixes: lane j's is row j's)
    for (int j = 0; j < nq; j++)
      q += (uint32_t) __builtin_amdgcn_readlane((int) oincl, j) <= k ? 1 : 0;
    q = q < 15 ? q : 15;
    const uint32_t base = (uint32_t) __shfl((int) (oincl - ocnt), q, 16);
    const uint32_t start = (uint32_t) __shfl((int) o0, q, 16);
    // own row q is halo row oy + q: its first slot and LDS record
    const int hr = oy + q < 15 ? oy + q : 15;
    const uint32_t hs_r = (uint32_t) __shfl((int) a0, hr, 16);
    const uint32_t lb_r = (uint32_t) __shfl((int) (incl - cnt), hr, 16);
    if (k < own_total) {
      const uint32_t gi = start + (k - base);
      own((int) k, gi, lb_r + (gi - hs_r));
    }
  }
  // halo row grp: slots [hs, hs + n), records [lb, lb + n)
  const uint32_t hs = (uint32_t) __shfl((int) a0, grp & 15, 16);
  const uint32_t lb = (uint32_t) __shfl((int) (incl - cnt), grp & 15, 16);
  const uint32_t n = (uint32_t) __shfl((int) cnt, grp & 15, 16);
  if (!c.ovf && grp < nr) {
    for (uint32_t k0 = (uint32_t) lane16; k0 < n; k0 += kBatch * 16) {
      F rr[kBatch][NS];
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const uint32_t k = k0 + (uint32_t) (u * 16);
        if (k < n) {
          if constexpr (policy_relative<M>::value) M::load(a, hs + k, rr[u], c.org);
          else M::load(a, hs + k, rr[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const uint32_t k = k0 + (uint32_t) (u * 16);
        if (k < n) {
#pragma unroll
          for (int v = 0; v < NS; v++) c.rec[(size_t) NS * (lb + k) + v] = rr[u][v];
        }
      }
    }
  }
  // the LDS tables (group 0 holds the same prefixes as every group)
  TT &T = *c.T;
  if (eok) T.lcell[threadIdx.x] = raw;
  for (int e = (int) threadIdx.x + (int) blockDim.x; e < ne; e += blockDim.x) {
    if (e % LW < w) T.lcell[e] = TT::halo_start(g, G, e / LW, e % LW);
  }
  if (threadIdx.x < 16u) {
    if (lane16 < nr) {
      T.hstart[lane16] = a0;
      T.lbase[lane16 + 1] = incl;
    }
    if (lane16 < nq) {
      T.ownrow[lane16] = (uint32_t) (oy + lane16);
      T.ownbase[lane16 + 1] = oincl;
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      T.t0[d] = G.T0[d];
      T.tn[d] = G.TN[d];
      T.h0[d] = G.H0[d];
      T.hn[d] = G.HN[d];
    }
    T.nrows = nr;
    T.lbase[0] = 0;
    T.ownbase[0] = 0;
    T.total = (int) total;
    T.own_total = (int) own_total;
    T.overflow = c.ovf ? 1 : 0;
    // record `total`: a position no near test accepts, the list padding
    if (!c.ovf) {
#pragma unroll
      for (int v = 0; v < NS; v++)
        c.rec[(size_t) NS * total + v] = v < DIM ? (F) kFarAway<F> : (F) 0;
    }
  }
  (void) ox;
  __syncthreads();
  return c;
}

template <class M, int TX, int TY, int TZ, int RM>
__device__ __forceinline__ TileCtx<M, TX, TY, TZ, RM> tile_stage(const typename M::Args &a,
                                                                 unsigned tile,
                                                                 unsigned char *smem) {
  return tile_stage<M, TX, TY, TZ, RM>(a, tile, smem, [](const TileCtx<M, TX, TY, TZ, RM> &) {});
}

template <class M, class Ctx>
__device__ __forceinline__ void tile_load(const typename M::Args &a, const Ctx &c, uint32_t j,
                                          typename M::F *r) {
  if constexpr (policy_relative<M>::value) M::load(a, j, r, c.org);
  else M::load(a, j, r);
}

// The own agent at global slot gi (LDS record ol): its record and FP64
// position; OVF: the records are in global memory (the halo did not fit LDS).
template <class M, bool OVF, class Ctx>
__device__ __forceinline__ void tile_own_at(const typename M::Args &a, const Ctx &c, uint32_t gi,
                                            uint32_t ol, typename M::F *own, double *pd) {
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  if constexpr (!OVF) {
#pragma unroll
    for (int v = 0; v < NS; v++) own[v] = c.rec[(size_t) NS * ol + v];
  } else {
    tile_load<M>(a, c, gi, own);
  }
  // the agent's cell comes from its stored position, exactly as binned
  pd[0] = pd[1] = pd[2] = 0.0;
  if constexpr (policy_relative<M>::value) {
    M::position(a, gi, pd);
  } else {
#pragma unroll
    for (int d = 0; d < DIM; d++) pd[d] = (double) own[d];
  }
}

// Own agent k of the tile (global slot gi, LDS record ol): record and FP64
// position; returns gi.
template <class M, bool OVF, class Ctx>
__device__ __forceinline__ uint32_t tile_own_src(const typename M::Args &a, const Ctx &c, int k,
                                                 typename M::F *own, double *pd) {
  int orow;
  uint32_t ol;
  c.T->own(k, &orow, &ol);
  const uint32_t gi = c.T->global_of(orow, ol);
  tile_own_at<M, OVF>(a, c, gi, ol, own, pd);
  return gi;
}

template <class M, class Ctx>
__device__ __forceinline__ uint32_t tile_own(const typename M::Args &a, const Ctx &c, int k,
                                             typename M::F *own, double *pd) {
  if (!c.ovf) return tile_own_src<M, false>(a, c, k, own, pd);
  return tile_own_src<M, true>(a, c, k, own, pd);
}

// The agent's cell (from its stored position, exactly as binned) and its
// fractional position inside the cell.
template <int DIM>
__device__ __forceinline__ int3 own_cell(const GridView &g, const double *pd, float *u) {
  const int3 cell = cell_of<double>(g, pd[0], pd[1], pd[2]);
  u[0] = (float) ((pd[0] - g.lo[0]) * g.inv_cell - (double) cell.x);
  u[1] = (float) ((pd[1] - g.lo[1]) * g.inv_cell - (double) cell.y);
  u[2] = DIM == 3 ? (float) ((pd[2] - g.lo[2]) * g.inv_cell - (double) cell.z) : 0.f;
  return cell;
}

// Walk the (narrowed) stencil of radius rc cell units around an own agent:
// fn(nb, l) for every candidate record nb, l its halo index, row by row of
// the stencil, each row in cell-sorted order.
template <class M, bool OVF, class Ctx, class Fn>
__device__ __forceinline__ void tile_walk_src(const typename M::Args &a, const Ctx &c, int3 cell,
                                              const float *u, float rc, Fn &fn) {
  using F = typename M::F;
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  const auto &T = *c.T;
  const int R = a.g.reach;
  const int zr = DIM == 3 ? R : 0;
  for (int dz = -zr; dz <= zr; dz++) {
    const int zz = cell.z + dz;
    if (zz < T.h0[2] || zz >= T.h0[2] + T.hn[2]) continue;
    for (int dy = -R; dy <= R; dy++) {
      const int yy = cell.y + dy;
      if (yy < T.h0[1] || yy >= T.h0[1] + T.hn[1]) continue;
      int xa, xb;
      if (!row_interval(u[0], u[1], u[2], dy, dz, rc, R, &xa, &xb)) continue;
      const int xlo = max(cell.x + xa, T.h0[0]);
      const int xhi = min(cell.x + xb, T.h0[0] + T.hn[0] - 1);
      if (xlo > xhi) continue;
      const int r = (zz - T.h0[2]) * T.hn[1] + (yy - T.h0[1]);
      const uint32_t l0 = T.cell_lds(r, xlo - T.h0[0]);
      const uint32_t l1 = T.cell_lds(r, xhi - T.h0[0] + 1);
      if constexpr (!OVF) {
        if (l0 < l1) {
          // two records in flight: the next one is read from LDS while the
          // current one is visited (ra and rb alternate, no copies)
          F ra[NS], rb[NS];
#pragma unroll
          for (int v = 0; v < NS; v++) ra[v] = c.rec[(size_t) NS * l0 + v];
          for (uint32_t l = l0;; l += 2) {
            const bool hb = l + 1 < l1;
            const uint32_t lb = hb ? l + 1 : l;
#pragma unroll
            for (int v = 0; v < NS; v++) rb[v] = c.rec[(size_t) NS * lb + v];
            fn(ra, l);
            if (!hb) break;
            const bool ha = l + 2 < l1;
            const uint32_t la = ha ? l + 2 : l + 1;
#pragma unroll
            for (int v = 0; v < NS; v++) ra[v] = c.rec[(size_t) NS * la + v];
            fn(rb, l + 1);
            if (!ha) break;
          }
        }
      } else {
        const uint32_t g0 = T.global_of(r, l0);
        for (uint32_t j = g0; j < g0 + (l1 - l0); j++) {
          F nb[NS];
          tile_load<M>(a, c, j, nb);
          fn(nb, l0 + (j - g0));
        }
      }
    }
  }
}

// The tile's records are in LDS or (overflow) in global memory for the whole
// block: branch once, so each loop body is specialised for its source.
template <class M, class Ctx, class Fn>
__device__ __forceinline__ void tile_walk(const typename M::Args &a, const Ctx &c, int3 cell,
                                          const float *u, float rc, Fn &&fn) {
  if (!c.ovf) tile_walk_src<M, false>(a, c, cell, u, rc, fn);
  else tile_walk_src<M, true>(a, c, cell, u, rc, fn);
}

// Entry e of agent i's list.
__device__ __forceinline__ size_t list_entry(const NListArgs &nl, uint32_t i, uint32_t e) {
  return ((size_t) (e >> 3) * (size_t) nl.stride + i) * 8u + (e & 7u);
}

// Visit the records of one agent's neighbour list: eight entries per 16-byte
// row, the next row loaded while this one is visited.  The caller supplies
// row 0 (staged with the tile).
template <class M, bool OVF, class Ctx, class Fn>
__device__ __forceinline__ void list_walk(const typename M::Args &a, const Ctx &c,
                                          const NListArgs &nl, uint32_t gi, uint32_t len,
                                          uint4 row0, Fn &fn) {
  using F = typename M::F;
  constexpr int NS = M::NS;
  // row r of this agent: four words at lp + r * rstride
  const uint32_t *lp = reinterpret_cast<const uint32_t *>(nl.idx) + (size_t) gi * 4u;
  const size_t rstride = (size_t) nl.stride * 4u;
  // four words (eight entries) per row, held as scalars
  uint32_t c0 = row0.x, c1 = row0.y, c2 = row0.z, c3 = row0.w;
  if constexpr (!OVF) {
    // lists are padded to whole rows of eight with the far-away record.
    // Two records in flight, as in tile_walk_src: the record of the next
    // entry is read from LDS before the current one is visited, so the
    // visit's branches never wait on an LDS read.
    F ra[NS], rb[NS];
    auto rd = [&](F *r, uint32_t l) {
#pragma unroll
      for (int v = 0; v < NS; v++) r[v] = c.rec[(size_t) NS * l + v];
    };
    // rows are read two ahead: row k + 1 is in m0..m3 while row k is
    // visited and row k + 2 is in flight
    uint32_t m0, m1, m2, m3;
    {
      const uint32_t *np = lp + (len > 8 ? rstride : 0);
      m0 = np[0];
      m1 = np[1];
      m2 = np[2];
      m3 = np[3];
    }
    rd(ra, c0 & 0xFFFFu);
    for (uint32_t e = 0; e < len; e += 8) {
      const uint32_t *np = lp + (size_t) ((e + 16 < len ? e + 16 : e) >> 3) * rstride;
      const uint32_t n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
      rd(rb, c0 >> 16);
      fn(ra);
      rd(ra, c1 & 0xFFFFu);
      fn(rb);
      rd(rb, c1 >> 16);
      fn(ra);
      rd(ra, c2 & 0xFFFFu);
      fn(rb);
      rd(rb, c2 >> 16);
      fn(ra);
      rd(ra, c3 & 0xFFFFu);
      fn(rb);
      rd(rb, c3 >> 16);
      fn(ra);
      // first entry of the next row (a re-read after the last row)
      rd(ra, m0 & 0xFFFFu);
      fn(rb);
      c0 = m0;
      c1 = m1;
      c2 = m2;
      c3 = m3;
      m0 = n0;
      m1 = n1;
      m2 = n2;
      m3 = n3;
    }
  } else {
    auto one = [&](uint32_t l) {
      if (l < (uint32_t) c.T->total) {
        F nb[NS];
        tile_load<M>(a, c, c.T->global_of(c.T->row_of_record(l), l), nb);
        fn(nb);
      }
    };
    for (uint32_t e = 0; e < len; e += 8) {
      const bool more = e + 8 < len;
      const uint32_t *np = lp + (size_t) ((more ? e + 8 : e) >> 3) * rstride;
      const uint32_t n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
      one(c0 & 0xFFFFu);
      one(c0 >> 16);
      one(c1 & 0xFFFFu);
      one(c1 >> 16);
      one(c2 & 0xFFFFu);
      one(c2 >> 16);
      one(c3 & 0xFFFFu);
      one(c3 >> 16);
      c0 = n0;
      c1 = n1;
      c2 = n2;
      c3 = n3;
    }
  }
}

template <class M, int TX, int TY, int TZ, int RM>
__device__ __forceinline__ void tiled_body(const typename M::Args &a, unsigned tile,
                                           unsigned char *smem) {
  using F = typename M::F;
  constexpr int NS = M::NS;
  const auto c = tile_stage<M, TX, TY, TZ, RM>(a, tile, smem);
  const int own_total = uniform(c.T->own_total);
  // the agent loop is specialised on where the records are (LDS or, for an
  // overflowing halo, global memory), so each form addresses one memory
  auto agents = [&](auto src) {
    constexpr bool OVF = decltype(src)::value;
    for (int k = threadIdx.x; k < own_total; k += blockDim.x) {
      F own[NS];
      double pd[3];
      float u[3];
      const uint32_t gi = tile_own_src<M, OVF>(a, c, k, own, pd);
      const int3 cell = own_cell<M::DIM>(a.g, pd, u);
      if constexpr (policy_agent<M>::value) {
        // the agent's step with its near loop over the stencil's records
        auto walk = [&](auto &&visit) {
          bool go = true;
          auto fn = [&](const F *nb, uint32_t) {
            if (go) go = visit(nb);
          };
          tile_walk_src<M, OVF>(a, c, cell, u, a.rc, fn);
        };
        double np[3];
        M::agent(a, gi, walk, np);
      } else {
        typename M::Acc acc;
        if constexpr (policy_relative<M>::value) M::begin(a, acc, own, pd, c.org);
        else M::begin(a, acc, own);
        auto fn = [&](const F *nb, uint32_t) { acc.visit(a, own, nb); };
        tile_walk_src<M, OVF>(a, c, cell, u, a.rc, fn);
        if constexpr (policy_relative<M>::value) M::finish(a, acc, own, gi, pd, nullptr);
        else M::finish(a, acc, own, gi, nullptr);
      }
    }
  };
  if (c.ovf) agents(std::true_type{});
  else agents(std::false_type{});
}

// ---- neighbour lists -------------------------------------------------------------
// Per-agent lists of the halo records within radius + skin, built by the
// stencil walk right after a (re-)sort and reused while the stored order is
// left alone: as long as no agent has moved more than skin / 2 since the
// build, every pair within the radius now was within radius + skin then, so
// walking the list (and applying the exact near test) finds exactly the
// stencil's neighbours, in the stencil's order of the build.
// Entries are halo indices of the build's tile (uint16), in rows of eight
// interleaved over the agents (list_entry) so a wave reads them coalesced.
// The step kernel flags a rebuild (gate of the next timestep) when an agent
// strays past skin / 2 or a list overflowed; the rebuild kernels of the next
// timestep run only when their gate is set.
// Build the lists of one tile (after a sort; records come from the sorted
// buffers a points at).
// Build the lists of one tile's agents; returns the longest list.
template <class M, int TX, int TY, int TZ, int RM>
__device__ __forceinline__ uint32_t list_build_tile(const typename M::Args &a, const NListArgs &nl,
                                                    unsigned tile, unsigned char *smem) {
  using F = typename M::F;
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  const auto c = tile_stage<M, TX, TY, TZ, RM>(a, tile, smem);
  uint32_t longest = 0;
  const int own_total = uniform(c.T->own_total);
  // list entries and the lane order hold 16-bit halo and own indices: a tile
  // with more records keeps no lists (its agents walk the stencil)
  const bool tile_lists = uniform(c.T->total) < (int) kListOverflow;
  for (int k = threadIdx.x; k < own_total; k += blockDim.x) {
    F own[NS];
    double pd[3];
    float u[3];
    const uint32_t gi = tile_own<M>(a, c, k, own, pd);
    const int3 cell = own_cell<M::DIM>(a.g, pd, u);
    uint32_t cnt = 0;
    // the current row of eight entries, stored as one 16-byte write when full
    uint64_t lo = 0, hi = 0;
    auto put = [&](uint32_t e, uint32_t l) {
      const uint64_t v = (uint64_t) (l & 0xFFFFu) << (16u * (e & 3u));
      if (e < 4u) lo |= v;
      else hi |= v;
    };
    auto store_row = [&](uint32_t r) {
      reinterpret_cast<uint4 *>(nl.idx)[(size_t) r * (size_t) nl.stride + gi] =
          make_uint4((uint32_t) lo, (uint32_t) (lo >> 32), (uint32_t) hi, (uint32_t) (hi >> 32));
      lo = hi = 0;
    };
    tile_walk<M>(a, c, cell, u, nl.rc_list, [&](const F *nb, uint32_t l) {
      double d2 = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double e = (double) nb[d] - (double) own[d];
        d2 += e * e;
      }
      if (d2 <= nl.r2_list) {
        if (cnt < (uint32_t) nl.cap) {
          put(cnt & 7u, l);
          if ((cnt & 7u) == 7u) store_row(cnt >> 3);
        }
        cnt++;
      }
    });
    const bool fits = tile_lists && cnt <= (uint32_t) nl.cap && cnt < kListOverflow;
    nl.len[gi] = fits ? (uint16_t) cnt : kListOverflow;
    if (fits && (cnt & 7u)) {
      // pad the last row with the far-away record (index total)
      for (uint32_t e = cnt & 7u; e < 8u; e++) put(e, (uint32_t) c.T->total);
      store_row(cnt >> 3);
    }
    if (tile_lists) longest = max(longest, cnt);
#pragma unroll
    for (int d = 0; d < DIM; d++) nl.pos0[(size_t) d * nl.n + gi] = pd[d];
    if (nl.copy.ncols) {
      gather_cols(nl.copy, gi, gi);
      nl.id_dst[gi] = nl.id_src[gi];
      nl.skey_dst[gi] = nl.skey_src[gi];
    }
  }
  // the order the step kernel's lanes take the tile's agents in: by list rows
  // (a counting sort; inside a bucket the order is arbitrary, and which lane
  // takes an agent never changes its results), stored at the slots of the
  // own agents in own order with each agent's length and first list row
  constexpr uint32_t kBuckets = 16;  // rows 0..14 (capped), 15 = overflowed
  __shared__ uint32_t bucket[kBuckets];
  auto slot_of = [&](int k) {
    int orow;
    uint32_t ol;
    c.T->own(k, &orow, &ol);
    return c.T->global_of(orow, ol);
  };
  auto bucket_of = [](uint32_t len) {
    return len == kListOverflow ? kBuckets - 1 : min((len + 7u) >> 3, kBuckets - 2);
  };
  if (!tile_lists) {
    for (int k = threadIdx.x; k < own_total; k += blockDim.x) nl.perm[slot_of(k)] = kOwnOrder;
    return longest;
  }
  if (threadIdx.x < kBuckets) bucket[threadIdx.x] = 0u;
  __syncthreads();  // also: every agent's length and list are written
  for (int k = threadIdx.x; k < own_total; k += blockDim.x)
    atomicAdd(&bucket[bucket_of(nl.len[slot_of(k)])], 1u);
  __syncthreads();
  if (threadIdx.x < (unsigned) kWave) {
    const uint32_t v = threadIdx.x < kBuckets ? bucket[threadIdx.x] : 0u;
    const uint32_t incl = row16_incl_scan(v);
    if (threadIdx.x < kBuckets) bucket[threadIdx.x] = incl - v;
  }
  __syncthreads();
  const uint4 *rows = reinterpret_cast<const uint4 *>(nl.idx);
  for (int k = threadIdx.x; k < own_total; k += blockDim.x) {
    const uint32_t gi = slot_of(k);
    const uint32_t len = nl.len[gi];
    const uint32_t pos = atomicAdd(&bucket[bucket_of(len)], 1u);
    const uint32_t gp = slot_of((int) pos);
    nl.perm[gp] = ((uint32_t) k << 16) | len;
    nl.prow0[gp] = rows[gi];
  }
  return longest;
}

// The new cell offsets become the current ones (grid-strided share) after a
// gated re-sort.
__device__ __forceinline__ void list_copy_starts(const NListArgs &nl) {
  if (!nl.start_dst) return;
  for (int k = (int) (blockIdx.x * blockDim.x + threadIdx.x); k <= nl.ncells;
       k += (int) (gridDim.x * blockDim.x))
    nl.start_dst[k] = nl.start_src[k];
}

// One atomic per block for the longest list.
__device__ __forceinline__ void list_longest(const NListArgs &nl, uint32_t longest) {
  __shared__ uint32_t blk_longest;
  if (threadIdx.x == 0) blk_longest = 0u;
  __syncthreads();
  if (longest) atomicMax(&blk_longest, longest);
  __syncthreads();
  if (threadIdx.x == 0 && blk_longest) atomicMax(&nl.stats[0], blk_longest);
}

template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256) void tiled_list_build(typename M::Args a, NListArgs nl) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (nl.gated && nl.gate[nl.parity] == 0u) return;
  if (nl.gated && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&nl.stats[2], 1u);  // rebuilds run
  list_copy_starts(nl);
  const uint32_t longest =
      list_build_tile<M, TX, TY, TZ, RM>(a, nl, xcd_block(blockIdx.x, gridDim.x), smem);
  list_longest(nl, longest);
}

// Positions as stored: FP64 for FP32-compensated policies, else the policy's type.
template <class M>
using policy_pos_t = typename std::conditional<policy_relative<M>::value, double, typename M::F>::type;

// The gated re-sort and list build of one timestep in one launch (256
// threads): bin, scan, place (abl_pipeline.hpp), then the copy of the new
// cell offsets and every tile's lists, as phases of a work queue whose waits
// end whether or not the blocks are co-resident (abl_pipeline.hpp).  Returns
// at once when the previous step set no gate, or after a wait ran past its
// spin limit (the host then sees q->error and rebuilds from scratch).
template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256) void tiled_rebuild(typename M::Args a, NListArgs nl,
                                                      DeltaArgs d) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (nl.gate[nl.parity] == 0u) return;
  if (__hip_atomic_load(&d.q->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&nl.stats[2], 1u);  // rebuilds run
  __shared__ uint32_t lds[kWave + 1];
  const unsigned nt = tile_count<TX, TY, TZ, M::DIM>(a.g);
  const uint32_t chunks = (uint32_t) (d.n + 255) / 256u;
  const uint32_t ncopy = nl.start_dst ? (uint32_t) (nl.ncells + 1 + 255) / 256u : 0u;
  QueuePlan plan;
  plan.end[0] = chunks;
  plan.end[1] = plan.end[0] + (uint32_t) d.nscan;
  plan.end[2] = plan.end[1] + chunks;
  plan.end[3] = plan.end[2] + ncopy + nt;
  uint32_t longest = 0;
  for (;;) {
    const uint32_t item = queue_claim(d.q);
    if (item >= plan.end[3]) break;
    const int phase = plan.phase_of(item);
    if (phase > 0 && !queue_wait(d.q, phase - 1, plan.items(phase - 1), d.spins)) break;
    const uint32_t k = item - plan.first(phase);
    if (phase < 3) {
      delta_item<policy_pos_t<M>>(d, phase, k, lds);
    } else if (k < ncopy) {
      const int c = (int) (k * 256u + threadIdx.x);
      if (c <= nl.ncells) nl.start_dst[c] = nl.start_src[c];
    } else {
      const unsigned t = k - ncopy;
      longest = max(longest, list_build_tile<M, TX, TY, TZ, RM>(a, nl, xcd_block(t, nt), smem));
    }
    queue_finish(d.q, phase);
  }
  list_longest(nl, longest);
  queue_exit(d.q);
}

// One timestep of the step through the lists.  An agent whose list
// overflowed walks the stencil (the grid is fresh: an overflow always
// forces a rebuild on the next timestep).
//
// The tile's agents are taken blockDim at a time, in the order the list
// build stored (by number of list rows), so each wave walks lists of about
// one length instead of every wave waiting for its longest.  Each lane's
// agent, list length and first list row are read together with the tile's
// records (one round of global reads).  Which lane handles an agent never
// changes the agent's results.
//
// Returns whether one of the thread's agents moved more than skin / 2 since
// the build or walked the stencil (a rebuild is due).
template <class M, int TX, int TY, int TZ, int RM>
__device__ __forceinline__ bool step_list_tile(const typename M::Args &a, const NListArgs &nl,
                                               unsigned tile, unsigned char *smem) {
  using F = typename M::F;
  using Ctx = TileCtx<M, TX, TY, TZ, RM>;
  using TT = typename Ctx::TT;
  constexpr int DIM = M::DIM;
  constexpr int NS = M::NS;
  // own agent k (< 256) of the row-group staging: its slot and LDS record
  __shared__ uint2 ownmap[256];
  // the agent of position k of the tile's order (own index << 16 | list
  // length) and its first list row, stored by the build at own slot k
  uint32_t my_perm = 0;
  uint4 my_row = make_uint4(0u, 0u, 0u, 0u);
  auto fetch = [&](const Ctx &cc, int k) {
    if (k < cc.T->own_total) {
      int orow;
      uint32_t ol;
      cc.T->own(k, &orow, &ol);
      const uint32_t gi = cc.T->global_of(orow, ol);
      my_perm = nl.perm[gi];
      my_row = nl.prow0[gi];
    }
  };
  // row-group staging (which fills ownmap) when the tile shape allows it
  const bool mapped = DIM == 2 && TT::MAXROWS <= 16 && blockDim.x == 256u;
  auto stage = [&]() {
    if constexpr (DIM == 2 && TT::MAXROWS <= 16) {
      if (mapped)
        return tile_stage_rows<M, TX, TY, TZ, RM>(a, tile, smem,
                                                  [&](int k, uint32_t gi, uint32_t ol) {
                                                    ownmap[k] = make_uint2(gi, ol);
                                                    my_perm = nl.perm[gi];
                                                    my_row = nl.prow0[gi];
                                                  });
    }
    return tile_stage<M, TX, TY, TZ, RM>(a, tile, smem,
                                         [&](const Ctx &cc) { fetch(cc, (int) threadIdx.x); });
  };
  auto c = stage();
  // the LDS pointers again from smem: through the two staging paths the
  // compiler loses that they point to LDS and would address it with flat
  // instructions
  c.T = reinterpret_cast<TT *>(smem);
  c.rec = reinterpret_cast<F *>(smem + ((sizeof(TT) + 15) & ~(size_t) 15));
  const int own_total = uniform(c.T->own_total);
  bool stray = false;
  for (int k0 = 0; k0 < own_total; k0 += (int) blockDim.x) {
    if (k0 > 0) fetch(c, k0 + (int) threadIdx.x);
    // lanes take the pass's positions of the tile's order: agents of about
    // one list length per wave
    if (k0 + (int) threadIdx.x >= own_total) continue;
    const bool own_order = my_perm == kOwnOrder;
    const int k = own_order ? k0 + (int) threadIdx.x : (int) (my_perm >> 16);
    const uint32_t len = own_order ? (uint32_t) kListOverflow : my_perm & 0xFFFFu;
    const uint4 row0 = my_row;

    // the whole agent body is specialised on where the records are, so the
    // LDS path never waits on global reads of the overflow path
    auto agent = [&](auto src) {
      constexpr bool OVF = decltype(src)::value;
      F own[NS];
      double pd[3];
      uint32_t gi;
      if (mapped && k < 256) {
        const uint2 m = ownmap[k];
        gi = m.x;
        tile_own_at<M, OVF>(a, c, gi, m.y, own, pd);
      } else {
        gi = tile_own_src<M, OVF>(a, c, k, own, pd);
      }
      double p0[3] = {0.0, 0.0, 0.0};
      double np[3] = {0.0, 0.0, 0.0};
      if constexpr (policy_agent<M>::value) {
        if (len == kListOverflow) stray = true;
        // the agent's step with its near loop over the list (or, for an
        // overflowed list, the stencil); a visit that returns false ends it
        auto walk = [&](auto &&visit) {
          bool go = true;
          if (len != kListOverflow) {
            auto fn = [&](const F *nb) {
              if (go) go = visit(nb);
            };
            list_walk<M, OVF>(a, c, nl, gi, len, row0, fn);
          } else {
            float u[3];
            const int3 cell = own_cell<M::DIM>(a.g, pd, u);
            auto fn = [&](const F *nb, uint32_t) {
              if (go) go = visit(nb);
            };
            tile_walk_src<M, OVF>(a, c, cell, u, a.rc, fn);
          }
        };
#pragma unroll
        for (int d = 0; d < DIM; d++) p0[d] = nl.pos0[(size_t) d * nl.n + gi];
        M::agent(a, gi, walk, np);
      } else {
        typename M::Acc acc;
        if constexpr (policy_relative<M>::value) M::begin(a, acc, own, pd, c.org);
        else M::begin(a, acc, own);
        if (len != kListOverflow) {
          auto fn = [&](const F *nb) { acc.visit(a, own, nb); };
          list_walk<M, OVF>(a, c, nl, gi, len, row0, fn);
        } else {
          stray = true;
          float u[3];
          const int3 cell = own_cell<M::DIM>(a.g, pd, u);
          auto fn = [&](const F *nb, uint32_t) { acc.visit(a, own, nb); };
          tile_walk_src<M, OVF>(a, c, cell, u, a.rc, fn);
        }
        // the build's position, read before the step's stores so its
        // latency overlaps the epilogue
#pragma unroll
        for (int d = 0; d < DIM; d++) p0[d] = nl.pos0[(size_t) d * nl.n + gi];
        if constexpr (policy_relative<M>::value) M::finish(a, acc, own, gi, pd, np);
        else M::finish(a, acc, own, gi, np);
      }
      double d2 = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double e = np[d] - p0[d];
        d2 += e * e;
      }
      return d2;
    };
    const double d2 = c.ovf ? agent(std::true_type{}) : agent(std::false_type{});
    // !(d2 <= bound) also catches NaN positions
    if (!(d2 <= nl.half_skin2)) stray = true;
  }
  return stray;
}

template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256, policy_list_waves<M>::value) void tiled_step_list(
    typename M::Args a, NListArgs nl) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (blockIdx.x == 0 && threadIdx.x == 0) nl.gate[nl.parity] = 0u;
  if (step_list_tile<M, TX, TY, TZ, RM>(a, nl, xcd_block(blockIdx.x, gridDim.x), smem))
    nl.gate[nl.parity ^ 1] = 1u;
}

// ---- ensemble forms of the list build and the list step ------------------------
// The tiles of many members in one launch (member lookup as in
// tiled_step_multi) over one set of lists indexed by the members'
// concatenated slots.  Besides the ensemble's gate (nl.gate: some member is
// due a rebuild) every member has its own, by timestep parity: mgate[p *
// members + m].  A member rebuilds exactly when a run of it alone would: when
// one of its own agents strayed or overflowed.

__device__ __forceinline__ int tile_member(const uint32_t *__restrict__ tile_base, int members,
                                           unsigned t) {
  int lo = 0, hi = members - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_base[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The lists of the tiles of members due a rebuild, after the gated sort
// (records in the sorted buffers args point at; copied back as in
// tiled_list_build).
template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256) void tiled_list_build_multi(const typename M::Args *__restrict__ args,
                                                              const uint32_t *__restrict__ tile_base,
                                                              int members, NListArgs nl,
                                                              const uint32_t *__restrict__ mgate) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (nl.gate[nl.parity] == 0u) return;
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int m = tile_member(tile_base, members, t);
  if (mgate[nl.parity * members + m] == 0u) return;
  const uint32_t longest = list_build_tile<M, TX, TY, TZ, RM>(args[m], nl, t - tile_base[m], smem);
  list_longest(nl, longest);
}

// One timestep of every member through the lists; clears this parity's gates
// and sets the next parity's for the members that strayed.
template <class M, int TX, int TY, int TZ, int RM>
__global__ __launch_bounds__(256, policy_list_waves<M>::value) void tiled_step_list_multi(
    const typename M::Args *__restrict__ args, const uint32_t *__restrict__ tile_base,
    int members, NListArgs nl, uint32_t *__restrict__ mgate) {
  extern __shared__ __align__(16) unsigned char smem[];
  const unsigned t = xcd_block(blockIdx.x, gridDim.x);
  const int m = tile_member(tile_base, members, t);
  const unsigned first = tile_base[m];
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0) nl.gate[nl.parity] = 0u;
    if (t == first) mgate[nl.parity * members + m] = 0u;
  }
  if (step_list_tile<M, TX, TY, TZ, RM>(args[m], nl, t - first, smem)) {
    nl.gate[nl.parity ^ 1] = 1u;
    mgate[(nl.parity ^ 1) * members + m] = 1u;
  }
}

// Host side: LDS record capacity (`slack` times the expected halo population
// plus 64, within a 64 KiB block budget less the 2 KiB of static LDS the list
// step kernel holds) and the block's dynamic LDS bytes.
// Less slack leaves room for more resident blocks per CU; a tile whose halo
// does not fit reads from global memory instead (same results).
template <class M, int TX, int TY, int TZ, int RM>
inline void tile_budget(float occupancy, int *cap, size_t *smem, float slack = 2.0f) {
  using F = typename M::F;
  using TT = TileTables<M::DIM, TX, TY, TZ, RM>;
  const size_t head = (sizeof(TT) + 15) & ~(size_t) 15;
  const double halo_cells = (double) (TX + 2 * RM) * (TY + 2 * RM) *
                            (M::DIM == 3 ? (double) (TZ + 2 * RM) : 1.0);
  const size_t budget = 62 * 1024 - head;
  const int max_cap = (int) (budget / (M::NS * sizeof(F)));
  const int want = (int) ((double) slack * halo_cells * (double) occupancy) + 64;
  *cap = want < max_cap ? want : max_cap;
  *smem = head + (size_t) *cap * M::NS * sizeof(F);
}

// Launch tiled_step for a tile shape.
template <class M, int TX, int TY, int TZ, int RM>
inline void launch_tiled(typename M::Args a, const GridView &g, float occupancy, int block,
                         hipStream_t stream, float slack = 2.0f) {
  size_t smem = 0;
  tile_budget<M, TX, TY, TZ, RM>(occupancy, &a.cap, &smem, slack);
  const int threads = block > 256 ? 256 : block;
  const unsigned ntiles = tile_count<TX, TY, TZ, M::DIM>(g);
  hipLaunchKernelGGL((tiled_step<M, TX, TY, TZ, RM>), ntiles, threads, smem, stream, a);
}

// Launch the neighbour-list build (mode 1) or the step through the lists
// (mode 2) for a tile shape; same tiles and LDS budget as tiled_step.
template <class M, int TX, int TY, int TZ, int RM>
inline void launch_list(typename M::Args a, const GridView &g, float occupancy, int block,
                        hipStream_t stream, float slack, const NListArgs &nl, int mode,
                        const DeltaArgs *delta) {
  size_t smem = 0;
  tile_budget<M, TX, TY, TZ, RM>(occupancy, &a.cap, &smem, slack);
  const int threads = block > 256 ? 256 : block;
  const unsigned ntiles = tile_count<TX, TY, TZ, M::DIM>(g);
  if (mode == 3) {
    // as many blocks as the CUs hold at once (a multiple of the 8 XCDs); the
    // work queue does not need them resident together, more would only
    // find the queue empty
    if (!delta || delta->pos_elem != (int) sizeof(policy_pos_t<M>) || !delta->launched) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    // the grid for (device, dynamic LDS) of the last launch, kept: the
    // occupancy query costs more host time than the launch
    static std::atomic<uint64_t> cached{0};
    const uint64_t key = ((uint64_t) (uint32_t) dev << 40) | ((uint64_t) smem << 16);
    uint64_t got = cached.load(std::memory_order_relaxed);
    unsigned grid = (got & ~0xFFFFull) == key ? (unsigned) (got & 0xFFFFu) : 0u;
    if (!grid) {
      int per_cu = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &per_cu, reinterpret_cast<const void *>(&tiled_rebuild<M, TX, TY, TZ, RM>), 256,
              smem) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return;
      grid = std::min(((unsigned) per_cu * (unsigned) cus) & ~7u, 0xFFF8u);
      if (grid < 8u) return;
      cached.store(key | grid, std::memory_order_relaxed);
    }
    if (delta->blocks) grid = delta->blocks;
    hipLaunchKernelGGL((tiled_rebuild<M, TX, TY, TZ, RM>), grid, 256, smem, stream, a, nl, *delta);
    *delta->launched = 1;
    return;
  }
  if (mode == 1) {
    hipLaunchKernelGGL((tiled_list_build<M, TX, TY, TZ, RM>), ntiles, threads, smem, stream, a, nl);
  } else {
    hipLaunchKernelGGL((tiled_step_list<M, TX, TY, TZ, RM>), ntiles, threads, smem, stream, a, nl);
  }
}

// Dispatch on the stencil reach (RM = 1, 2 or 4 cells) and on the list mode.
template <class M, int TX, int TY, int TZ>
inline void launch_tiles(const StepLaunch &L, const typename M::Args &a) {
  const int R = L.grid.reach;
  if (L.nl_mode) {
    if (R <= 1) launch_list<M, TX, TY, TZ, 1>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack, *L.nl, L.nl_mode, L.delta);
    else if (R == 2) launch_list<M, TX, TY, TZ, 2>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack, *L.nl, L.nl_mode, L.delta);
    else launch_list<M, TX, TY, TZ, 4>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack, *L.nl, L.nl_mode, L.delta);
  } else {
    if (R <= 1) launch_tiled<M, TX, TY, TZ, 1>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack);
    else if (R == 2) launch_tiled<M, TX, TY, TZ, 2>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack);
    else launch_tiled<M, TX, TY, TZ, 4>(a, L.grid, L.occupancy, L.block, L.stream, L.lds_slack);
  }
}

}  // namespace abl
