// Deterministic, load-balanced row sums over a CSR/CSC: sum_{e in row r}
// value(e, col[e]) for every row, the pull step of PageRank (over the
// transpose) and of its adjoint (over the CSR).
//
// Layout (merge path).  A sequential CSR walk visits n row ends and m edges
// in one interleaved order; that sequence is cut into tiles of kTile items,
// so every tile holds exactly kTile row ends + edges whatever the degree mix
// (empty rows, hubs and everything between cost the same per item).  The
// plan stores each tile's first row and first edge.
//
// Tile work (one 512-lane workgroup, two resident per CU).
//  1. Stage: a software pipeline keeps the next tile's gathers, row ends and
//     finalisation operands, and the columns of the tile after it, in
//     flight while a tile is processed (consecutive lanes on consecutive
//     edges; values of the hottest sources come from an LDS cache).  A
//     tile's values and row ends are then stored to LDS.
//  2. Walk: lane t takes merge-path items [t*kUnits, (t+1)*kUnits), finds its
//     starting (row, edge) by binary search in LDS and walks them in order,
//     adding edge values and emitting a row sum at every row end.
//  3. Carries: the partial of the row left open at the end of each lane is
//     combined by a segmented scan over lanes (fixed tree order); the lane
//     that emits a row adds the carries of the earlier lanes of that row.
//  4. Finalise: the tile's row sums are written to LDS and finalised by
//     consecutive lanes, so the per-row loads and stores are coalesced.
// A row that began in an earlier tile is not finalised in the tile where it
// ends: that tile stores its partial (head) and every tile stores the partial
// of the row open at its end (tail); a second small launch adds the tails of
// the earlier tiles to the head, in tile order, and finalises those rows.
// Sums are compensated (Neumaier) in FP32 and plain in FP64; every reduction
// has a fixed order, so results are bit-reproducible run to run.
//
// Src  : __device__ Real operator()(int64_t e, int32_t c) const  (value of
//        edge slot e whose column id is c)
// Fin  : __device__ void load()  (read per-sweep scalars, once per thread)
//        struct Pre; __device__ Pre pre(uint32_t r) const  (per-row operands,
//        loaded early so their latency overlaps the tile's gathers)
//        __device__ void operator()(uint32_t r, Real sum, const Pre&,
//        double& a0, double& a1) const  (finalise row r; two FP64
//        per-workgroup accumulators, e.g. residual and dangling mass)
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>

#include "gx_device.hpp"
#include "gx_internal.hpp"

namespace gx {
namespace rs {

constexpr int kBlock = 256;                 // generic kernels (reductions, spanning rows)
constexpr int kSweepBlock = 512;            // tile workgroup: 8 wavefronts, two per CU with the LDS cache
constexpr int kUnits = 8;                   // merge-path items per lane
constexpr int kTile = kSweepBlock * kUnits; // items (row ends + edges) per tile
constexpr int kWaves = kSweepBlock / kWave;
constexpr int kSweepGrid = 256 * 8;         // upper bound of the persistent tile launch
constexpr int kFixGrid = 256;               // workgroups of the spanning-row launch
constexpr int kParts = kSweepGrid + kFixGrid;
constexpr int kItems = kTile / kSweepBlock; // edges per lane at most (a tile has <= kTile edges)
// LDS cache of the gathered vector for sources [0, kHot) (the hottest ones
// in degree order), sized to what the tile leaves of the 160 KB of LDS.
template <class Real>
struct HotCache {
  static constexpr int kHot = sizeof(Real) == 4 ? 11264 : 1024;
};

template <class Real>
struct Acc;

template <>
struct Acc<float> {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(float os, float oc) {
    add(os);
    c += oc;
  }
  __device__ __forceinline__ float value() const { return s + c; }
  __device__ __forceinline__ float comp() const { return c; }
};

template <>
struct Acc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double x) { s += x; }
  __device__ __forceinline__ void merge(double os, double) { s += os; }
  __device__ __forceinline__ double value() const { return s; }
  __device__ __forceinline__ double comp() const { return 0.0; }
};

// Workgroup sum of one double per lane in a fixed order (valid in lane 0);
// NT = threads of the workgroup, lds holds NT / 64 doubles.
template <int NT = kBlock>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const unsigned w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / kWave; ++i) r += lds[i];
  __syncthreads();
  return r;
}

template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
typedef int v4i __attribute__((ext_vector_type(4)));

struct RowPlan {
  const int32_t* rowoff;  // n+1
  const int32_t* col;     // len
  int64_t n;              // rows
  int64_t len;            // edges
  int32_t* tile_row;      // ntiles+1: row of each tile's first item
  int32_t* tile_edge;     // ntiles+1: edge of each tile's first item
  int64_t ntiles;
};

// Builds (or returns the cached) plan of a row-offset array; defined in
// rowsweep.hip.
RowPlan build_plan(gx_graph* g, const int32_t* rowoff, const int32_t* col, int64_t n, int64_t m, PullPlan& cache);

// LDS index with one pad word per kUnits: lanes walk items kUnits apart, and
// the pad spreads them over the banks.
__device__ __forceinline__ int pad8(int j) { return j + j / kUnits; }

template <class Real>
struct TileShared {
  Real val[kTile + kTile / kUnits];          // edge values (padded index)
  alignas(16) char rows[kTile * sizeof(Real)];  // relative row ends, then row sums
  Real ws[kWaves], wc[kWaves];               // per-wave tail segment of the carries
  int wkey[kWaves], wfirst[kWaves];          // its key; key of the wave's lane 0
  Real ts[kWaves], tc[kWaves];               // tile tail partial per wave
};

// Inclusive scan of (s, c) over the lanes of a wavefront, restarted wherever
// `key` changes (equal keys are contiguous).  Fixed tree order.
template <class Real>
__device__ __forceinline__ Acc<Real> wave_seg_scan(int key, Acc<Real> a) {
  const int lane = static_cast<int>(lane_id());
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const Real os = __shfl_up(a.s, off);
    const Real oc = __shfl_up(a.comp(), off);
    const int ok = __shfl_up(key, off);
    if (lane >= off && ok == key) a.merge(os, oc);
  }
  return a;
}

struct TileCoord {
  int32_t r0, r1, e0, e1;
  bool defer0;  // row r0 began in an earlier tile and ends in this one
};

__device__ __forceinline__ TileCoord tile_coord(const RowPlan& P, int64_t T) {
  TileCoord c;
  c.r0 = P.tile_row[T];
  c.r1 = P.tile_row[T + 1];
  c.e0 = P.tile_edge[T];
  c.e1 = P.tile_edge[T + 1];
  c.defer0 = c.r1 > c.r0 && P.rowoff[c.r0] < c.e0;
  return c;
}

// Columns of a tile: edge e0 + tid + i*kSweepBlock in cc[i].  Consecutive lanes
// hold consecutive edges, so the gathers of one wavefront instruction hit
// runs of neighbouring sources (rows in degree order have ascending, and for
// heavy rows dense, source ids) that the texture unit coalesces into few
// cache lines.
__device__ __forceinline__ void tile_cols(const RowPlan& P, const TileCoord& tc, int32_t (&cc)[kItems]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t q = static_cast<int64_t>(tc.e0) + tid + i * kSweepBlock;
    cc[i] = q < tc.e1 ? ld_stream(&P.col[q]) : 0;
  }
}

// The loads of one tile, issued ahead of the tile's processing: columns
// (two tiles ahead), then gathered values, the lane's first row end and the
// finalisation operands of the lane's first row (one tile ahead).
template <class Real, class Fin>
struct Stage {
  TileCoord tc;
  bool valid;
  int32_t cc[kItems];
  Real v[kItems];
  int32_t rend;
  typename Fin::Pre pre;
};

template <class Real, class Fin>
__device__ __forceinline__ void stage_cols(const RowPlan& P, int64_t T, Stage<Real, Fin>& st) {
  st.valid = T < P.ntiles;
  if (st.valid) {
    st.tc = tile_coord(P, T);
    tile_cols(P, st.tc, st.cc);
  }
}

template <class Real, int HOT, class Src, class Fin>
__device__ __forceinline__ void stage_gather(const RowPlan& P, const Src& src, const Fin& fin, const Real* hot,
                                             Stage<Real, Fin>& st) {
  if (!st.valid) return;
  const int tid = threadIdx.x;
  const int32_t e0 = st.tc.e0, e1 = st.tc.e1;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t q = static_cast<int64_t>(e0) + tid + i * kSweepBlock;
    if constexpr (HOT > 0)
      st.v[i] = q < e1 ? src.apply(q, st.cc[i] < HOT ? hot[st.cc[i]] : src.load(st.cc[i])) : Real(0);
    else
      st.v[i] = q < e1 ? src(q, st.cc[i]) : Real(0);
  }
  const int nre = st.tc.r1 - st.tc.r0;
  st.rend = tid < nre ? ld_stream(&P.rowoff[st.tc.r0 + 1 + tid]) : 0;
  st.pre = tid < nre ? fin.pre(static_cast<uint32_t>(st.tc.r0 + tid)) : typename Fin::Pre{};
}

// One tile (see the file comment); its loads were issued by stage_gather
// one tile earlier.
template <class Real, class Src, class Fin>
__device__ __forceinline__ void sweep_tile(int64_t T, const RowPlan& P, const Fin& fin, TileShared<Real>& sh,
                                           Real* __restrict__ carry, const Stage<Real, Fin>& st, double& a0,
                                           double& a1) {
  const int tid = threadIdx.x;
  const int lane = static_cast<int>(lane_id());
  const int wid = tid / kWave;
  const TileCoord& tc = st.tc;
  const int32_t r0 = tc.r0, e0 = tc.e0, e1 = tc.e1;
  const int nre = tc.r1 - r0;  // row ends in this tile
  const int ned = e1 - e0;     // edges in this tile
  const int L = nre + ned;
  const bool defer0 = tc.defer0;
  int32_t* s_end = reinterpret_cast<int32_t*>(sh.rows);
  Real* s_sum = reinterpret_cast<Real*>(sh.rows);
  const typename Fin::Pre& pre0 = st.pre;

  // 1. Values and row ends to LDS.
  if (tid < nre) s_end[tid] = st.rend - e0;
  for (int j = tid + kSweepBlock; j < nre; j += kSweepBlock) s_end[j] = ld_stream(&P.rowoff[r0 + 1 + j]) - e0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int j = tid + i * kSweepBlock;
    if (j < ned) sh.val[pad8(j)] = st.v[i];
  }
  __syncthreads();

  // 2. Walk this lane's merge-path items.
  const int d = tid * kUnits;
  int rr = 0, ee = 0;
  if (d < L) {
    int lo = max(0, d - ned), hi = min(d, nre);  // rows ended before item d
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_end[mid - 1] + mid <= d) lo = mid; else hi = mid - 1;
    }
    rr = lo;
    ee = d - lo;
  }
  Acc<Real> acc, first;
  Real em[kUnits];
  uint32_t emask = 0;
  int rstart = 0;
#pragma unroll
  for (int u = 0; u < kUnits; ++u) {
    em[u] = Real(0);
    if (d + u < L) {
      if (rr < nre && ee >= s_end[rr]) {
        if (emask == 0) {
          first = acc;
          rstart = rr;
        }
        em[u] = acc.value();
        emask |= 1u << u;
        acc = Acc<Real>();
        ++rr;
      } else {
        acc.add(sh.val[pad8(ee)]);
        ++ee;
      }
    }
  }

  // 3. Carries: key = the row left open by this lane.
  const int key = (d < L) ? rr : INT_MAX;
  const Acc<Real> seg = wave_seg_scan(key, acc);
  const Real prev_s = __shfl_up(seg.s, 1);
  const Real prev_c = __shfl_up(seg.comp(), 1);
  const int prev_key = __shfl_up(key, 1);
  const int key0 = __shfl(key, 0);
  if (lane == kWave - 1) {
    sh.ws[wid] = seg.s;
    sh.wc[wid] = seg.comp();
    sh.wkey[wid] = key;
  }
  if (lane == 0) sh.wfirst[wid] = key;
  // Tile tail: the carries of the row open at the tile's end (local nre).
  Acc<Real> tail;
  if (key == nre) tail = acc;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const Real os = __shfl_xor(tail.s, off);
    const Real oc = __shfl_xor(tail.comp(), off);
    tail.merge(os, oc);
  }
  if (lane == 0) {
    sh.ts[wid] = tail.s;
    sh.tc[wid] = tail.comp();
  }
  __syncthreads();  // walk done: s_end may now be overwritten by row sums

  if (emask) {
    Acc<Real> f = first;
    if (lane > 0 && prev_key == rstart) f.merge(prev_s, prev_c);
    if (key0 == rstart || lane == 0) {
      // The run of this row reaches lane 0: add earlier wavefronts whose
      // trailing segment belongs to it.
      for (int w = wid - 1; w >= 0 && sh.wkey[w] == rstart; --w) {
        f.merge(sh.ws[w], sh.wc[w]);
        if (sh.wfirst[w] != rstart) break;
      }
    }
    const Real fval = f.value();
    int k = 0;
#pragma unroll
    for (int u = 0; u < kUnits; ++u) {
      if ((emask >> u) & 1u) {
        const int r = rstart + k;
        const Real val = (k == 0) ? fval : em[u];
        if (r == 0 && defer0)
          carry[T] = val;  // head of a row that began in an earlier tile
        else
          s_sum[r] = val;
        ++k;
      }
    }
  }
  if (tid == 0) {
    Acc<Real> t;
    for (int w = 0; w < kWaves; ++w) t.merge(sh.ts[w], sh.tc[w]);
    carry[P.ntiles + T] = t.s;
    carry[2 * P.ntiles + T] = t.comp();
  }
  __syncthreads();

  // 4. Finalise the rows that end here (coalesced).
  if (tid < nre && (tid > 0 || !defer0)) fin(static_cast<uint32_t>(r0 + tid), s_sum[tid], pre0, a0, a1);
  for (int j = tid + kSweepBlock; j < nre; j += kSweepBlock)
    fin(static_cast<uint32_t>(r0 + j), s_sum[j], fin.pre(static_cast<uint32_t>(r0 + j)), a0, a1);
  __syncthreads();
}

// One launch over every tile (persistent workgroups; tile T runs on
// workgroup T mod gridDim.x).  Workgroup partials of the two accumulators
// go to part0/part1[blockIdx.x].  Returns at once when *skip is non-zero (a
// converged iteration loop).  With HOT > 0 the gathered vector of sources
// [0, HOT) is first copied to LDS (src.load(c) is the gathered value,
// src.apply(e, value) the edge's term).
template <class Real, int HOT, class Src, class Fin>
__global__ void __launch_bounds__(kSweepBlock) sweep_kernel(RowPlan P, Src src, Fin fin,
                                                            const unsigned* __restrict__ skip,
                                                            double* __restrict__ part0, double* __restrict__ part1,
                                                            Real* __restrict__ carry) {
  if (skip && *skip) return;
  fin.load();
  __shared__ TileShared<Real> sh;
  __shared__ double red[kWaves];
  __shared__ Real hot[HOT > 0 ? HOT : 1];
  if constexpr (HOT > 0) {
    const int64_t nh = P.n < HOT ? P.n : HOT;
    for (int64_t i = threadIdx.x; i < nh; i += kSweepBlock) hot[i] = src.load(static_cast<int32_t>(i));
    __syncthreads();
  }
  double a0 = 0, a1 = 0;
  // Software pipeline: tile T is processed while the gathers of tile
  // T + grid and the columns of tile T + 2*grid are in flight.
  const int64_t G = gridDim.x;
  int64_t T = blockIdx.x;
  Stage<Real, Fin> cur, nxt;
  stage_cols(P, T, cur);
  stage_gather<Real, HOT>(P, src, fin, hot, cur);
  stage_cols(P, T + G, nxt);
  for (; T < P.ntiles; T += G) {
    stage_gather<Real, HOT>(P, src, fin, hot, nxt);
    Stage<Real, Fin> after;
    stage_cols(P, T + 2 * G, after);
    sweep_tile<Real, Src>(T, P, fin, sh, carry, cur, a0, a1);
    cur = nxt;
    nxt = after;
  }
  a0 = block_sum<kSweepBlock>(a0, red);
  a1 = block_sum<kSweepBlock>(a1, red);
  if (threadIdx.x == 0) {
    part0[blockIdx.x] = a0;
    part1[blockIdx.x] = a1;
  }
}

// Rows that span tiles: head of the tile where the row ends plus the tails
// of the earlier tiles, latest first.
template <class Real, class Fin>
__global__ void __launch_bounds__(kBlock) span_finalize_kernel(RowPlan P, Fin fin, const unsigned* __restrict__ skip,
                                                               const Real* __restrict__ carry,
                                                               double* __restrict__ part0,
                                                               double* __restrict__ part1) {
  if (skip && *skip) return;
  fin.load();
  __shared__ double red[kBlock / kWave];
  double a0 = 0, a1 = 0;
  for (int64_t T = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; T < P.ntiles;
       T += static_cast<int64_t>(gridDim.x) * kBlock) {
    const int32_t r0 = P.tile_row[T];
    if (r0 >= P.tile_row[T + 1] || P.rowoff[r0] >= P.tile_edge[T]) continue;
    Acc<Real> a;
    a.merge(carry[T], Real(0));
    for (int64_t t2 = T - 1; t2 >= 0 && P.tile_row[t2 + 1] == r0; --t2)
      a.merge(carry[P.ntiles + t2], carry[2 * P.ntiles + t2]);
    fin(static_cast<uint32_t>(r0), a.value(), fin.pre(static_cast<uint32_t>(r0)), a0, a1);
  }
  a0 = block_sum(a0, red);
  a1 = block_sum(a1, red);
  if (threadIdx.x == 0) {
    part0[blockIdx.x] = a0;
    part1[blockIdx.x] = a1;
  }
}

// Host: one full sweep (tile launch + spanning-row launch) on stream s.
// part0/part1 hold kParts entries each and carry 3 * P.ntiles; returns the
// number of partials written (at most kParts, contiguous from 0).  `timer`
// brackets the tile launch.  HOT > 0 caches the gathered values of sources
// [0, HOT) in LDS (only worth it when the lowest ids are the most gathered,
// as in the degree-ordered PageRank pull).
template <class Real, int HOT = 0, class Src, class Fin>
int launch_sweep(const RowPlan& P, const Src& src, const Fin& fin, const unsigned* skip, double* part0,
                 double* part1, Real* carry, hipStream_t s, KernelTimer* timer) {
  // Persistent grid: every workgroup resident at once (occupancy x CUs).
  static const int grid = [] {
    int per_cu = 0, dev = 0, cus = 0;
    GX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_kernel<Real, HOT, Src, Fin>, kSweepBlock,
                                                        0));
    GX_HIP(hipGetDevice(&dev));
    GX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return std::max(1, std::min(kSweepGrid, per_cu * cus));
  }();
  if (timer) timer->begin();
  sweep_kernel<Real, HOT, Src, Fin><<<grid, kSweepBlock, 0, s>>>(P, src, fin, skip, part0, part1, carry);
  if (timer) timer->end();
  GX_LAUNCHED("sweep_kernel");
  span_finalize_kernel<Real, Fin><<<kFixGrid, kBlock, 0, s>>>(P, fin, skip, carry, part0 + grid, part1 + grid);
  GX_LAUNCHED("span_finalize_kernel");
  return grid + kFixGrid;
}

}  // namespace rs
}  // namespace gx
