// Deterministic, load-balanced row sums over a CSR/CSC: sum_{e in row r}
// value(e, col[e]) for every row, the pull step of PageRank (over the
// transpose) and of its adjoint (over the CSR).
//
// Layout.  The edge array is cut into windows of kWin edges; each row is
// assigned to the window that holds its first edge, so a window owns a run of
// consecutive rows and at most 2*kWin contiguous edges.  Rows longer than
// kWin are cut into pieces of kPiece edges.  One launch covers every window
// and every piece (items interleaved over the grid); a second small launch
// combines the piece partials of the long rows in piece order.
//
// Window work.  Lane t of a 256-lane workgroup owns 16 consecutive edge
// slots (16-byte column loads), evaluates their values into registers and
// sums them row segment by row segment.  A row inside one lane's slots is
// finalised by that lane; a row that crosses lanes is finalised by the lane
// holding its first edge, which adds the continuation partials left in LDS by
// the later lanes, in lane order.  Sums are compensated (Neumaier) in FP32,
// plain in FP64; every reduction has a fixed order, so results are
// bit-reproducible run to run.
//
// Src  : __device__ Real operator()(int64_t e, int32_t c) const  (value of
//        edge slot e whose column id is c)
// Fin  : __device__ void load()  (read per-sweep scalars, once per thread)
//        __device__ void operator()(uint32_t r, Real sum, double& a0,
//        double& a1) const  (finalise row r; two FP64 per-workgroup
//        accumulators, e.g. residual and dangling mass)
#pragma once

#include <cmath>

#include "gx_device.hpp"
#include "gx_internal.hpp"

namespace gx {
namespace rs {

constexpr int kBlock = 256;
constexpr int kWin = 2040;       // edges per window (rows up to kWin stay in windows)
constexpr int kPiece = 4096;     // edges per long-row piece
constexpr int kPerLane = 16;     // edge slots per lane
constexpr int kRowCap = 1024;    // rows per window held in LDS
constexpr int kSweepGrid = 256 * 8;
constexpr int kLongGrid = 2048;
static_assert(kPerLane * kBlock >= 2 * kWin + 3, "window must fit the lanes' slots");

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

template <class Real>
__device__ __forceinline__ Acc<Real> wave_reduce(Acc<Real> a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const Real os = __shfl_xor(a.s, off);
    const Real oc = __shfl_xor(a.comp(), off);
    a.merge(os, oc);
  }
  return a;
}

// Workgroup sum of one double per lane in a fixed order (valid in lane 0).
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const unsigned w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / kWave; ++i) r += lds[i];
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
  int64_t len;            // number of edges
  int32_t* win_rows;      // nwin+1 first rows of windows (last = n)
  int64_t nwin;
  uint32_t* long_rows;    // rows with more than kWin edges
  uint32_t* piece_first;  // nlong+1 exclusive prefix of pieces per long row
  uint32_t* piece_row;    // owning long-row index of each piece
  int64_t nlong, npieces;
};

// Builds (or returns the cached) plan of a row-offset array; defined in
// rowsweep.hip.
RowPlan build_plan(gx_graph* g, const int32_t* rowoff, const int32_t* col, int64_t n, int64_t m, PullPlan& cache);

template <class Real>
struct SweepShared {
  int32_t off[kRowCap + 1];
  Real cs[kBlock];
  Real cc[kBlock];
};

template <class Real, class Src, class Fin>
__device__ __forceinline__ void sweep_window(int64_t k, const RowPlan& P, const Src& src, const Fin& fin,
                                             SweepShared<Real>& sh, double& a0, double& a1) {
  const int tid = threadIdx.x;
  const int32_t* __restrict__ rowoff = P.rowoff;
  const int32_t r0 = P.win_rows[k], r1 = P.win_rows[k + 1];
  if (r0 == r1) return;
  int32_t rl = r1;
  if (rowoff[r1] - rowoff[r1 - 1] > kWin) rl = r1 - 1;  // long last row: piece items
  const int32_t nrows = rl - r0;
  if (nrows == 0) return;
  const int32_t e0 = rowoff[r0];
  const int32_t nnz = rowoff[rl] - e0;
  if (nrows > kRowCap) {
    // Rare (long runs of empty rows): a lane per row.
    for (int32_t r = r0 + tid; r < rl; r += kBlock) {
      Acc<Real> a;
      for (int32_t i = rowoff[r], end = rowoff[r + 1]; i < end; ++i) a.add(src(i, P.col[i]));
      fin(r, a.value(), a0, a1);
    }
    return;
  }
  for (int32_t i = tid; i <= nrows; i += kBlock) sh.off[i] = ld_stream(&rowoff[r0 + i]) - e0;
  __syncthreads();
  for (int32_t i = tid; i < nrows; i += kBlock)
    if (sh.off[i] == sh.off[i + 1]) fin(r0 + i, Real(0), a0, a1);
  // This lane's slots, relative to e0: [lo, hi) within [0, nnz).
  const int32_t A = e0 & ~3;
  const int32_t base_rel = A - e0 + tid * kPerLane;  // may be negative for lane 0
  const int32_t lo = max(base_rel, 0), hi = min(base_rel + kPerLane, nnz);
  Real v[kPerLane];
  if (lo < hi) {
    const int64_t abs0 = static_cast<int64_t>(A) + tid * kPerLane;
    if (abs0 + kPerLane <= P.len) {
      const v4i* p = reinterpret_cast<const v4i*>(P.col + abs0);
#pragma unroll
      for (int q = 0; q < kPerLane / 4; ++q) {
        const v4i c = ld_stream(p + q);
        const int32_t cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int32_t i = base_rel + 4 * q + j;
          v[4 * q + j] = (i >= lo && i < hi) ? src(abs0 + 4 * q + j, cc[j]) : Real(0);
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < kPerLane; ++q) {
        const int32_t i = base_rel + q;
        v[q] = (i >= lo && i < hi) ? src(abs0 + q, P.col[abs0 + q]) : Real(0);
      }
    }
  }
  // Segmented pass (no barrier inside: lanes of a wavefront may diverge).
  bool owner_open = false;
  int ri = 0;
  Acc<Real> head;  // first segment when it continues a row begun in an earlier lane
  Acc<Real> own;   // segment of the row this lane opened and that continues
  if (lo < hi) {
    int a = 0, b = nrows - 1;  // last row with offset <= lo
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (sh.off[mid] <= lo) a = mid; else b = mid - 1;
    }
    ri = a;
    int32_t rend = sh.off[ri + 1];
    bool started_here = sh.off[ri] >= lo;
    Acc<Real> acc;
#pragma unroll
    for (int q = 0; q < kPerLane; ++q) {
      const int32_t i = base_rel + q;
      if (i < lo || i >= hi) continue;
      while (i >= rend) {
        if (started_here)
          fin(r0 + ri, acc.value(), a0, a1);
        else
          head = acc;
        acc = Acc<Real>();
        ++ri;
        while (sh.off[ri + 1] == sh.off[ri]) ++ri;  // skip empty rows
        rend = sh.off[ri + 1];
        started_here = true;
      }
      acc.add(v[q]);
    }
    if (!started_here) {
      head = acc;
    } else if (rend <= hi) {
      fin(r0 + ri, acc.value(), a0, a1);
    } else {
      owner_open = true;
      own = acc;
    }
  }
  sh.cs[tid] = head.s;
  sh.cc[tid] = head.comp();
  __syncthreads();
  if (owner_open) {
    const int32_t rend = sh.off[ri + 1];
    const int last_lane = (rend - 1 - (A - e0)) / kPerLane;
    for (int t2 = tid + 1; t2 <= last_lane; ++t2) own.merge(sh.cs[t2], sh.cc[t2]);
    fin(r0 + ri, own.value(), a0, a1);
  }
  __syncthreads();
}

template <class Real, class Src>
__device__ __forceinline__ void sweep_piece(int64_t p, const RowPlan& P, const Src& src, Real* __restrict__ piece_s,
                                            Real* __restrict__ piece_c, SweepShared<Real>& sh) {
  const unsigned lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t li = P.piece_row[p];
  const uint32_t r = P.long_rows[li];
  const int64_t local = p - P.piece_first[li];
  const int64_t b = P.rowoff[r] + local * kPiece;
  const int64_t e = min(b + static_cast<int64_t>(kPiece), static_cast<int64_t>(P.rowoff[r + 1]));
  Acc<Real> a;
  for (int64_t i = b + threadIdx.x; i < e; i += kBlock) a.add(src(i, ld_stream(&P.col[i])));
  a = wave_reduce(a);
  if (lane == 0) {
    sh.cs[wid] = a.s;
    sh.cc[wid] = a.comp();
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc<Real> t;
    for (int w = 0; w < kBlock / kWave; ++w) t.merge(sh.cs[w], sh.cc[w]);
    piece_s[p] = t.s;
    piece_c[p] = t.comp();
  }
  __syncthreads();
}

// One launch: every window and every long-row piece.  Workgroup partials of
// the two accumulators go to part0/part1[blockIdx.x].  Returns at once when
// *skip is non-zero (a converged iteration loop).
template <class Real, class Src, class Fin>
__global__ void __launch_bounds__(kBlock) sweep_kernel(RowPlan P, Src src, Fin fin, const unsigned* __restrict__ skip,
                                                       double* __restrict__ part0, double* __restrict__ part1,
                                                       Real* __restrict__ piece_s, Real* __restrict__ piece_c) {
  if (skip && *skip) return;
  fin.load();
  __shared__ SweepShared<Real> sh;
  __shared__ double red[kBlock / kWave];
  double a0 = 0, a1 = 0;
  const int64_t items = P.nwin + P.npieces;
  for (int64_t k = blockIdx.x; k < items; k += gridDim.x) {
    if (k < P.nwin)
      sweep_window<Real>(k, P, src, fin, sh, a0, a1);
    else
      sweep_piece<Real>(k - P.nwin, P, src, piece_s, piece_c, sh);
  }
  a0 = block_sum(a0, red);
  a1 = block_sum(a1, red);
  if (threadIdx.x == 0) {
    part0[blockIdx.x] = a0;
    part1[blockIdx.x] = a1;
  }
}

// Long rows: combine the piece partials in order and finalise.
template <class Real, class Fin>
__global__ void __launch_bounds__(kBlock) long_finalize_kernel(RowPlan P, Fin fin, const unsigned* __restrict__ skip,
                                                               const Real* __restrict__ piece_s,
                                                               const Real* __restrict__ piece_c,
                                                               double* __restrict__ part0,
                                                               double* __restrict__ part1) {
  if (skip && *skip) return;
  fin.load();
  __shared__ double red[kBlock / kWave];
  double a0 = 0, a1 = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; i < P.nlong;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    Acc<Real> a;
    for (uint32_t p = P.piece_first[i]; p < P.piece_first[i + 1]; ++p) a.merge(piece_s[p], piece_c[p]);
    fin(P.long_rows[i], a.value(), a0, a1);
  }
  a0 = block_sum(a0, red);
  a1 = block_sum(a1, red);
  if (threadIdx.x == 0) {
    part0[blockIdx.x] = a0;
    part1[blockIdx.x] = a1;
  }
}

// Host: one full sweep (window/piece launch + long-row launch) on stream s.
// Partials: part0/part1 hold kSweepGrid + kLongGrid entries; returns the
// number of partials written.  `timer` brackets the main launch.
template <class Real, class Src, class Fin>
int launch_sweep(const RowPlan& P, const Src& src, const Fin& fin, const unsigned* skip, double* part0,
                 double* part1, Real* piece_s, Real* piece_c, hipStream_t s, KernelTimer* timer) {
  if (timer) timer->begin();
  sweep_kernel<Real, Src, Fin><<<kSweepGrid, kBlock, 0, s>>>(P, src, fin, skip, part0, part1, piece_s, piece_c);
  if (timer) timer->end();
  GX_LAUNCHED("sweep_kernel");
  if (P.nlong == 0) return kSweepGrid;
  long_finalize_kernel<Real, Fin><<<kLongGrid, kBlock, 0, s>>>(P, fin, skip, piece_s, piece_c, part0 + kSweepGrid,
                                                              part1 + kSweepGrid);
  GX_LAUNCHED("long_finalize_kernel");
  return kSweepGrid + kLongGrid;
}

}  // namespace rs
}  // namespace gx
