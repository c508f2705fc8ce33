// PageRank: normalised pull (FP32-Kahan / FP64), the reference's own
// unnormalised pull variants, and the synchronous residual push.
//
// Reference
//  - pull: lonestar/analytics/cpu/pagerank/PageRank-pull.cpp:196-251
//    (computePRTopological: x'[v] = alpha*sum_{u->v} x[u]/nout[u] + (1-a)/N,
//    L1 change as the stop test) and :132-190 (computePRResidual), with
//    computeOutDeg :90-122 and the constants of PageRank-constants.h:32-52.
//  - push: PageRank-push.cpp:103-188 (syncPageRank) and the GPU app
//    lonestar/analytics/gpu/pagerank/pagerank.cu:24-236.
//
// Normalised semantics used by gx_pagerank_pull: x0 = 1/n;
//   x'[v] = (1-a)/n + a*(sum_{u->v} y[u] + D/n),  y[u] = x[u]/outdeg(u),
//   D = sum of x over dangling vertices; stop when sum|x'-x| < tol (FP64) or
//   after max_iter iterations.
//
// Design of the pull sweep (gfx950, HBM/latency bound, no MFMA)
//  - The transpose (CSC) is cut into edge windows of W edges.  Rows are
//    assigned to the window holding their first in-edge, so a window's rows
//    are consecutive vertex ids and its edges are one contiguous range of at
//    most 2W.  One workgroup per window streams tcol coalesced, gathers y into
//    LDS, then reduces rows out of LDS: a lane per row when the window holds
//    many rows, a wavefront per row (strided lanes + wave reduction) when it
//    holds few.  Rows longer than W are cut into pieces of P edges reduced by
//    whole workgroups; their per-piece partials are combined in piece order.
//    The window table and the piece table are built once per graph.
//  - FP32 mode: compensated (Neumaier) row sums in registers and
//    TwoSum-merged wave reductions; x and y stored in FP32; the residual and
//    dangling mass are accumulated in FP64 per workgroup and reduced in a
//    fixed order (deterministic).
//  - Iterations are enqueued in batches with no host round trip: each sweep
//    reads a device `converged` flag and returns at once when it is set, so
//    surplus launches after convergence are no-ops.
#include <algorithm>
#include <cmath>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr int kWin = 2040;          // edges per window (rows up to kWin stay in windows)
constexpr int kPiece = 8192;        // edges per long-row piece
constexpr int kSweepGrid = 256 * 8; // workgroups of the window sweep
constexpr int kLongGrid = 2048;

// ------------------------------------------------------------ accumulators
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
  __device__ __forceinline__ void merge(const Acc& o) {
    add(o.s);
    c += o.c;
  }
  __device__ __forceinline__ float value() const { return s + c; }
  __device__ __forceinline__ Acc shfl_xor(int off) const {
    Acc r;
    r.s = __shfl_xor(s, off);
    r.c = __shfl_xor(c, off);
    return r;
  }
};

template <>
struct Acc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double x) { s += x; }
  __device__ __forceinline__ void merge(const Acc& o) { s += o.s; }
  __device__ __forceinline__ double value() const { return s; }
  __device__ __forceinline__ Acc shfl_xor(int off) const {
    Acc r;
    r.s = __shfl_xor(s, off);
    return r;
  }
};

template <class Real>
__device__ __forceinline__ Acc<Real> wave_reduce(Acc<Real> a) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a.merge(a.shfl_xor(off));
  return a;
}

// Block-wide fixed-order sum of one double per thread (result valid in
// thread 0).
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

struct PrState {
  double resid;
  double dang;
  double base;       // (1-a)/n + a*D/n for the next sweep
  double alpha;
  double tol;
  double inv_n;
  unsigned converged;
  unsigned iters;
  unsigned final_buf;  // x buffer holding the latest iterate
  unsigned max_iter;
};

// Per-graph tables for the window sweep.
struct PrPlan {
  int32_t* win_rows;    // nwin+1 first rows of windows (last = n)
  int64_t nwin;
  uint32_t* long_rows;  // rows with more than kWin in-edges
  uint32_t* piece_first;// nlong+1 exclusive prefix of pieces per long row
  int64_t nlong, npieces;
  uint32_t* piece_row;  // owning long-row index of each piece
};

// ------------------------------------------------------------- preprocessing
__global__ void win_rows_kernel(const int32_t* __restrict__ trow, int64_t n, int64_t nwin,
                                int32_t* __restrict__ win_rows) {
  int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; k <= nwin; k += stride) {
    if (k == nwin) {
      win_rows[k] = static_cast<int32_t>(n);
      continue;
    }
    const int64_t target = k * kWin;
    int64_t lo = 0, hi = n;  // first r in [0,n) with trow[r] >= target, else n
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (trow[mid] >= target) hi = mid; else lo = mid + 1;
    }
    win_rows[k] = static_cast<int32_t>(lo);
  }
}

__global__ void long_flags_kernel(const int32_t* __restrict__ trow, int64_t n, uint32_t* __restrict__ flag) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) flag[v] = (trow[v + 1] - trow[v] > kWin) ? 1u : 0u;
}

__global__ void long_compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                    int64_t n, uint32_t* __restrict__ out) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride)
    if (flag[v]) out[pos[v]] = static_cast<uint32_t>(v);
}

__global__ void piece_count_kernel(const int32_t* __restrict__ trow, const uint32_t* __restrict__ long_rows,
                                   int64_t nlong, uint32_t* __restrict__ cnt) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < nlong) {
    const uint32_t r = long_rows[i];
    const int64_t d = trow[r + 1] - trow[r];
    cnt[i] = static_cast<uint32_t>((d + kPiece - 1) / kPiece);
  } else if (i == nlong) {
    cnt[i] = 0;
  }
}

__global__ void piece_row_kernel(const uint32_t* __restrict__ piece_first, int64_t nlong,
                                 uint32_t* __restrict__ piece_row) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < nlong; i += stride)
    for (uint32_t p = piece_first[i]; p < piece_first[i + 1]; ++p) piece_row[p] = static_cast<uint32_t>(i);
}

// ------------------------------------------------------------------ sweeps
template <class Real>
__device__ __forceinline__ void finalize_row(uint32_t r, Real sum, const PrState& S, const Real* __restrict__ x_old,
                                             Real* __restrict__ x_new, Real* __restrict__ y_new,
                                             const uint32_t* __restrict__ outdeg, double& resid, double& dang) {
  const Real xn = static_cast<Real>(S.base) + static_cast<Real>(S.alpha) * sum;
  resid += fabs(static_cast<double>(xn) - static_cast<double>(x_old[r]));
  x_new[r] = xn;
  const uint32_t d = outdeg[r];
  if (d) {
    y_new[r] = xn / static_cast<Real>(d);
  } else {
    y_new[r] = Real(0);
    dang += static_cast<double>(xn);
  }
}

// One workgroup per window.  Lane t owns the 16 consecutive edge slots
// [A + 16t, A + 16t + 16) (A = window start rounded down to 4 so the column
// loads are 16-byte vectors), gathers their y values into registers and sums
// them row segment by row segment.  A row that lies inside a lane's slots is
// finalised by that lane; a row that crosses lanes is finalised by the lane
// holding its first edge, which adds the continuation partials that the later
// lanes left in LDS, in lane order (deterministic).  Empty rows get x' = base.
// Windows with more rows than kRowCap fall back to a lane per row.
constexpr int kPerLane = 16;
constexpr int kRowCap = 1024;
static_assert(kPerLane * kBlock >= 2 * kWin + 3, "window must fit the lanes' slots");

template <class Real>
__global__ void __launch_bounds__(kBlock) pr_window_kernel(
    const int32_t* __restrict__ trow, const int32_t* __restrict__ tcol, const int32_t* __restrict__ win_rows,
    int64_t nwin, const Real* __restrict__ y, const Real* __restrict__ x_old, Real* __restrict__ x_new,
    Real* __restrict__ y_new, const uint32_t* __restrict__ outdeg, const PrState* __restrict__ state,
    double* __restrict__ part_resid, double* __restrict__ part_dang) {
  if (state->converged) return;
  const PrState S = *state;
  __shared__ int32_t s_off[kRowCap + 1];
  __shared__ Real s_cs[kBlock];
  __shared__ Real s_cc[kBlock];
  __shared__ double red[kBlock / kWave];
  double resid = 0, dang = 0;
  const int tid = threadIdx.x;
  for (int64_t k = blockIdx.x; k < nwin; k += gridDim.x) {
    const int32_t r0 = win_rows[k], r1 = win_rows[k + 1];
    if (r0 == r1) continue;
    int32_t rl = r1;
    if (trow[r1] - trow[r1 - 1] > kWin) rl = r1 - 1;  // long last row: piece sweep
    const int32_t nrows = rl - r0;
    if (nrows == 0) continue;
    const int32_t e0 = trow[r0];
    const int32_t nnz = trow[rl] - e0;
    if (nrows > kRowCap) {
      // Rare (long runs of empty rows): a lane per row.
      for (int32_t r = r0 + tid; r < rl; r += kBlock) {
        Acc<Real> a;
        for (int32_t i = trow[r], end = trow[r + 1]; i < end; ++i) a.add(y[tcol[i]]);
        finalize_row<Real>(r, a.value(), S, x_old, x_new, y_new, outdeg, resid, dang);
      }
      continue;
    }
    for (int32_t i = tid; i <= nrows; i += kBlock) s_off[i] = trow[r0 + i] - e0;
    __syncthreads();
    // Empty rows.
    for (int32_t i = tid; i < nrows; i += kBlock)
      if (s_off[i] == s_off[i + 1]) finalize_row<Real>(r0 + i, Real(0), S, x_old, x_new, y_new, outdeg, resid, dang);
    // This lane's slots, relative to e0: [lo, hi) within [0, nnz).
    const int32_t A = e0 & ~3;
    const int32_t base_rel = A - e0 + tid * kPerLane;  // may be negative for lane 0
    const int32_t lo = max(base_rel, 0), hi = min(base_rel + kPerLane, nnz);
    Real v[kPerLane];
    if (base_rel < nnz) {
      const int4* p = reinterpret_cast<const int4*>(tcol + A + tid * kPerLane);
#pragma unroll
      for (int q = 0; q < kPerLane / 4; ++q) {
        const int4 c = p[q];
        v[4 * q + 0] = y[c.x];
        v[4 * q + 1] = y[c.y];
        v[4 * q + 2] = y[c.z];
        v[4 * q + 3] = y[c.w];
      }
    }
    // Segmented pass over this lane's slots (no barrier inside: lanes of a
    // wavefront may diverge here).
    const bool active = lo < hi;
    bool owner_open = false;  // this lane holds the first edge of a row that continues
    int ri = 0;
    Acc<Real> cont;
    if (active) {
      // Row (local index) holding slot lo: last row with offset <= lo.
      int a = 0, b = nrows - 1;
      while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (s_off[mid] <= lo) a = mid; else b = mid - 1;
      }
      ri = a;
      int32_t rend = s_off[ri + 1];
      bool started_here = s_off[ri] >= lo;
      Acc<Real> acc;
#pragma unroll
      for (int q = 0; q < kPerLane; ++q) {
        const int32_t i = base_rel + q;
        if (i < lo || i >= hi) continue;
        while (i >= rend) {
          // Row ri ended before slot i.
          if (started_here)
            finalize_row<Real>(r0 + ri, acc.value(), S, x_old, x_new, y_new, outdeg, resid, dang);
          else
            cont = acc;
          acc = Acc<Real>();
          ++ri;
          while (s_off[ri + 1] == s_off[ri]) ++ri;  // skip empty rows
          rend = s_off[ri + 1];
          started_here = true;
        }
        acc.add(v[q]);
      }
      if (!started_here) {
        cont = acc;  // whole range inside a row that began in an earlier lane
      } else if (rend <= hi) {
        finalize_row<Real>(r0 + ri, acc.value(), S, x_old, x_new, y_new, outdeg, resid, dang);
      } else {
        owner_open = true;
        cont = acc;
      }
    }
    s_cs[tid] = cont.s;
    if constexpr (sizeof(Real) == 4) s_cc[tid] = cont.c;
    __syncthreads();
    if (owner_open) {
      // Add the continuation partials of the lanes covering the rest of the row.
      const int32_t rend = s_off[ri + 1];
      const int last_lane = (rend - 1 - (A - e0)) / kPerLane;
      for (int t2 = tid + 1; t2 <= last_lane; ++t2) {
        Acc<Real> o;
        o.s = s_cs[t2];
        if constexpr (sizeof(Real) == 4) o.c = s_cc[t2];
        cont.merge(o);
      }
      finalize_row<Real>(r0 + ri, cont.value(), S, x_old, x_new, y_new, outdeg, resid, dang);
    }
    __syncthreads();
  }
  resid = block_sum(resid, red);
  dang = block_sum(dang, red);
  if (tid == 0) {
    part_resid[blockIdx.x] = resid;
    part_dang[blockIdx.x] = dang;
  }
}

// Partial (compensated) sums of long-row pieces.
template <class Real>
__global__ void __launch_bounds__(kBlock) pr_piece_kernel(
    const int32_t* __restrict__ trow, const int32_t* __restrict__ tcol, const uint32_t* __restrict__ long_rows,
    const uint32_t* __restrict__ piece_first, const uint32_t* __restrict__ piece_row, int64_t npieces,
    const Real* __restrict__ y, const PrState* __restrict__ state, Real* __restrict__ piece_s,
    Real* __restrict__ piece_c) {
  if (state->converged) return;
  __shared__ Real red_s[kBlock / kWave];
  __shared__ Real red_c[kBlock / kWave];
  const unsigned lane = lane_id(), wid = threadIdx.x / kWave;
  for (int64_t p = blockIdx.x; p < npieces; p += gridDim.x) {
    const uint32_t li = piece_row[p];
    const uint32_t r = long_rows[li];
    const int64_t local = p - piece_first[li];
    const int64_t b = trow[r] + local * kPiece;
    const int64_t e = min(b + static_cast<int64_t>(kPiece), static_cast<int64_t>(trow[r + 1]));
    Acc<Real> a;
    for (int64_t i = b + threadIdx.x; i < e; i += kBlock) a.add(y[tcol[i]]);
    a = wave_reduce(a);
    if (lane == 0) {
      red_s[wid] = a.s;
      if constexpr (sizeof(Real) == 4) red_c[wid] = a.c; else red_c[wid] = Real(0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Acc<Real> t;
      for (int w = 0; w < kBlock / kWave; ++w) {
        Acc<Real> o;
        o.s = red_s[w];
        if constexpr (sizeof(Real) == 4) o.c = red_c[w];
        t.merge(o);
      }
      piece_s[p] = t.s;
      if constexpr (sizeof(Real) == 4) piece_c[p] = t.c; else piece_c[p] = Real(0);
    }
    __syncthreads();
  }
}

// Long rows: combine piece partials in order, then finalise.
template <class Real>
__global__ void __launch_bounds__(kBlock) pr_long_finalize_kernel(
    const uint32_t* __restrict__ long_rows, const uint32_t* __restrict__ piece_first, int64_t nlong,
    const Real* __restrict__ piece_s, const Real* __restrict__ piece_c, const Real* __restrict__ x_old,
    Real* __restrict__ x_new, Real* __restrict__ y_new, const uint32_t* __restrict__ outdeg,
    const PrState* __restrict__ state, double* __restrict__ part_resid, double* __restrict__ part_dang) {
  if (state->converged) return;
  const PrState S = *state;
  __shared__ double red[kBlock / kWave];
  double resid = 0, dang = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; i < nlong;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    Acc<Real> a;
    for (uint32_t p = piece_first[i]; p < piece_first[i + 1]; ++p) {
      Acc<Real> t;
      if constexpr (sizeof(Real) == 4) {
        t.s = piece_s[p];
        t.c = piece_c[p];
      } else {
        t.s = piece_s[p];
      }
      a.merge(t);
    }
    finalize_row<Real>(long_rows[i], a.value(), S, x_old, x_new, y_new, outdeg, resid, dang);
  }
  resid = block_sum(resid, red);
  dang = block_sum(dang, red);
  if (threadIdx.x == 0) {
    part_resid[blockIdx.x] = resid;
    part_dang[blockIdx.x] = dang;
  }
}

// Fixed-order reduction of the per-workgroup partials; sets the next base,
// the convergence flag and the buffer index of the newest iterate.
__global__ void __launch_bounds__(kBlock) pr_reduce_kernel(PrState* state, const double* __restrict__ part_resid,
                                                           const double* __restrict__ part_dang, int nparts,
                                                           unsigned newest_buf) {
  if (state->converged) return;
  __shared__ double red[kBlock / kWave];
  double r = 0, d = 0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    r += part_resid[i];
    d += part_dang[i];
  }
  r = block_sum(r, red);
  d = block_sum(d, red);
  if (threadIdx.x == 0) {
    state->resid = r;
    state->dang = d;
    state->base = (1.0 - state->alpha) * state->inv_n + state->alpha * d * state->inv_n;
    state->iters += 1;
    state->final_buf = newest_buf;
    if (r < state->tol || state->iters >= state->max_iter) state->converged = 1;
  }
}

// x0 = 1/n, y0 = x0/outdeg; dangling partials.
template <class Real>
__global__ void __launch_bounds__(kBlock) pr_init_kernel(Real* __restrict__ x, Real* __restrict__ y,
                                                         const uint32_t* __restrict__ outdeg, int64_t n,
                                                         double* __restrict__ part_resid,
                                                         double* __restrict__ part_dang) {
  __shared__ double red[kBlock / kWave];
  const Real x0 = Real(1) / static_cast<Real>(n);
  double dang = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * kBlock) {
    x[v] = x0;
    const uint32_t d = outdeg[v];
    y[v] = d ? x0 / static_cast<Real>(d) : Real(0);
    if (!d) dang += static_cast<double>(x0);
  }
  dang = block_sum(dang, red);
  if (threadIdx.x == 0) {
    part_resid[blockIdx.x] = 0.0;
    part_dang[blockIdx.x] = dang;
  }
}

__global__ void pr_state_init_kernel(PrState* s, double alpha, double tol, double inv_n, unsigned max_iter) {
  if (threadIdx.x == 0) {
    s->resid = 0;
    s->dang = 0;
    s->base = 0;
    s->alpha = alpha;
    s->tol = tol;
    s->inv_n = inv_n;
    s->converged = 0;
    s->iters = 0;
    s->final_buf = 0;
    s->max_iter = max_iter;
  }
}

// Dangling mass of the initial vector -> base; does not count an iteration.
__global__ void pr_init_reduce_kernel(PrState* state, const double* __restrict__ part_dang, int nparts) {
  __shared__ double red[kBlock / kWave];
  double d = 0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) d += part_dang[i];
  d = block_sum(d, red);
  if (threadIdx.x == 0) {
    state->dang = d;
    state->base = (1.0 - state->alpha) * state->inv_n + state->alpha * d * state->inv_n;
    if (state->max_iter == 0) state->converged = 1;
  }
}

// Builds (once per graph) the window and long-row piece tables in
// g->pull_plan.buf: win_rows[nwin+1] | long_rows[nlong] | piece_first[nlong+1]
// | piece_row[npieces].
PrPlan plan_for(gx_graph* g) {
  PullPlan& pp = g->pull_plan;
  const int64_t n = g->n, m = g->m;
  const int64_t nwin = m / kWin + 1;
  const int64_t max_long = m / kWin + 1;
  const int64_t max_pieces = m / kPiece + max_long + 1;
  auto view = [&]() {
    PrPlan P{};
    int32_t* base = pp.buf.as<int32_t>();
    P.nwin = nwin;
    P.win_rows = base;
    P.long_rows = reinterpret_cast<uint32_t*>(base + (nwin + 1));
    P.piece_first = P.long_rows + max_long;
    P.piece_row = P.piece_first + (max_long + 1);
    P.nlong = pp.nlong;
    P.npieces = pp.npieces;
    return P;
  };
  if (pp.ready) return view();
  build_transpose(g);
  hipStream_t s = g->stream;
  pp.buf.ensure(((nwin + 1) + max_long + (max_long + 1) + max_pieces) * sizeof(int32_t));
  pp.nlong = pp.npieces = 0;
  PrPlan P = view();
  win_rows_kernel<<<grid_for(nwin + 1, kBlock, 4096), kBlock, 0, s>>>(g->trow(), n, nwin, P.win_rows);
  GX_LAUNCHED("win_rows_kernel");
  DevBuf flag, pos;
  flag.ensure(std::max<int64_t>(n, 1) * 4);
  pos.ensure(std::max<int64_t>(n, 1) * 4);
  long_flags_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(g->trow(), n, flag.as<uint32_t>());
  GX_LAUNCHED("long_flags_kernel");
  exclusive_scan_u32(flag.as<uint32_t>(), pos.as<uint32_t>(), n, s);
  uint32_t lastpos = 0, lastflag = 0;
  read_bytes(g, &lastpos, pos.as<uint32_t>() + n - 1, 4);
  read_bytes(g, &lastflag, flag.as<uint32_t>() + n - 1, 4);
  const int64_t nlong = static_cast<int64_t>(lastpos) + lastflag;
  GX_REQUIRE(nlong <= max_long, "PageRank long-row table bound exceeded");
  int64_t npieces = 0;
  if (nlong > 0) {
    long_compact_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(flag.as<uint32_t>(), pos.as<uint32_t>(), n,
                                                                      P.long_rows);
    GX_LAUNCHED("long_compact_kernel");
    DevBuf cnt;
    cnt.ensure((nlong + 1) * 4);
    piece_count_kernel<<<grid_for(nlong + 1, kBlock), kBlock, 0, s>>>(g->trow(), P.long_rows, nlong,
                                                                      cnt.as<uint32_t>());
    GX_LAUNCHED("piece_count_kernel");
    exclusive_scan_u32(cnt.as<uint32_t>(), P.piece_first, nlong + 1, s);
    uint32_t np = 0;
    read_bytes(g, &np, P.piece_first + nlong, 4);
    npieces = np;
    GX_REQUIRE(npieces <= max_pieces, "PageRank piece table bound exceeded");
    piece_row_kernel<<<grid_for(nlong, kBlock, 4096), kBlock, 0, s>>>(P.piece_first, nlong, P.piece_row);
    GX_LAUNCHED("piece_row_kernel");
  }
  GX_HIP(hipStreamSynchronize(s));
  pp.nlong = nlong;
  pp.npieces = npieces;
  pp.ready = true;
  return view();
}

}  // namespace

template <class Real>
static void run_pull(gx_graph* g, double alpha, double tol, uint32_t max_iter, Real* pr, gx_stats* st) {
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  ensure_outdeg(g);
  PrPlan P = plan_for(g);
  const uint32_t* outdeg = g->outdeg_buf.as<uint32_t>();
  Real* xb = g->scratch<Real>(WS_V0, 2 * n);
  Real* yb = g->scratch<Real>(WS_V1, 2 * n);
  Real* x[2] = {xb, xb + n};
  Real* y[2] = {yb, yb + n};
  const int nparts = kSweepGrid + kLongGrid;
  double* parts = g->scratch<double>(WS_REDUCE, 2 * nparts);
  double* part_resid = parts;
  double* part_dang = parts + nparts;
  Real* piece_s = g->scratch<Real>(WS_V2, std::max<int64_t>(P.npieces, 1));
  Real* piece_c = g->scratch<Real>(WS_V3, std::max<int64_t>(P.npieces, 1));
  PrState* state = g->scratch<PrState>(WS_COUNTERS, 1);
  PrState* host = static_cast<PrState*>(g->host_pinned(sizeof(PrState)));

  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0;
  pr_state_init_kernel<<<1, 64, 0, s>>>(state, alpha, tol, 1.0 / static_cast<double>(n), max_iter);
  pr_init_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(x[0], y[0], outdeg, n, part_resid, part_dang);
  GX_LAUNCHED("pr_init_kernel");
  pr_init_reduce_kernel<<<1, kBlock, 0, s>>>(state, part_dang, kSweepGrid);
  GX_LAUNCHED("pr_init_reduce_kernel");
  launches += 3;
  uint32_t it = 0;
  const uint32_t batch = 8;
  int cur = 0;
  while (it < max_iter) {
    const uint32_t end = std::min(max_iter, it + batch);
    for (; it < end; ++it) {
      const int nxt = cur ^ 1;
      ktimer.begin();
      pr_window_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(g->trow(), g->tcol(), P.win_rows, P.nwin, y[cur], x[cur],
                                                           x[nxt], y[nxt], outdeg, state, part_resid, part_dang);
      ktimer.end();
      GX_LAUNCHED("pr_window_kernel");
      int np = kSweepGrid;
      if (P.npieces > 0) {
        pr_piece_kernel<Real><<<kLongGrid, kBlock, 0, s>>>(g->trow(), g->tcol(), P.long_rows, P.piece_first,
                                                           P.piece_row, P.npieces, y[cur], state, piece_s, piece_c);
        GX_LAUNCHED("pr_piece_kernel");
        pr_long_finalize_kernel<Real><<<kLongGrid, kBlock, 0, s>>>(
            P.long_rows, P.piece_first, P.nlong, piece_s, piece_c, x[cur], x[nxt], y[nxt], outdeg, state,
            part_resid + kSweepGrid, part_dang + kSweepGrid);
        GX_LAUNCHED("pr_long_finalize_kernel");
        np = nparts;
        launches += 2;
      }
      pr_reduce_kernel<<<1, kBlock, 0, s>>>(state, part_resid, part_dang, np, static_cast<unsigned>(nxt));
      GX_LAUNCHED("pr_reduce_kernel");
      launches += 2;
      cur = nxt;
    }
    GX_HIP(hipMemcpyAsync(host, state, sizeof(PrState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->converged) break;
  }
  GX_HIP(hipMemcpyAsync(host, state, sizeof(PrState), hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  GX_HIP(hipMemcpyAsync(pr, x[host->final_buf], n * sizeof(Real), hipMemcpyDeviceToDevice, s));
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = host->iters;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(g->m) * host->iters;
    st->items = static_cast<uint64_t>(n) * host->iters;
    st->residual = host->resid;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
}

extern "C" gx_status gx_pagerank_pull(gx_graph* g, double alpha, double tol, uint32_t max_iter, int precision,
                                      void* pr, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(pr, "pr");
  GX_REQUIRE(g->n > 0, "graph has no vertices");
  GX_REQUIRE(alpha >= 0 && alpha < 1, "alpha must be in [0,1)");
  GX_REQUIRE(precision == GX_PR_FP32 || precision == GX_PR_FP64, "precision must be GX_PR_FP32 or GX_PR_FP64");
  if (precision == GX_PR_FP32)
    run_pull<float>(g, alpha, tol, max_iter, static_cast<float*>(pr), st);
  else
    run_pull<double>(g, alpha, tol, max_iter, static_cast<double*>(pr), st);
  GX_API_END
}

// ------------------------------------------------- reference-semantics pull
namespace {

// Phase 1 of computePRResidual (PageRank-pull.cpp:140-158).
__global__ void __launch_bounds__(kBlock) ref_residual_delta_kernel(float* __restrict__ value,
                                                                    float* __restrict__ residual,
                                                                    float* __restrict__ delta,
                                                                    const uint32_t* __restrict__ nout, int64_t n,
                                                                    float alpha, float tol,
                                                                    unsigned* __restrict__ accum) {
#pragma clang fp contract(off)
  unsigned cnt = 0;
  for (int64_t s = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; s < n;
       s += static_cast<int64_t>(gridDim.x) * kBlock) {
    float d = 0.0f;
    const float r = residual[s];
    if (r > tol) {
      residual[s] = 0.0f;
      value[s] = value[s] + r;
      const uint32_t k = nout[s];
      if (k > 0) {
        d = (r * alpha) / static_cast<float>(k);
        ++cnt;
      }
    }
    delta[s] = d;
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(accum, cnt);
}

// Phase 2 (PageRank-pull.cpp:160-175): sequential FP32 sum over in-edges in
// CSC order, so the result is bit-identical to the CPU path.
__global__ void __launch_bounds__(kBlock) ref_residual_sum_kernel(const int32_t* __restrict__ trow,
                                                                  const int32_t* __restrict__ tcol,
                                                                  const float* __restrict__ delta,
                                                                  float* __restrict__ residual, int64_t n) {
#pragma clang fp contract(off)
  for (int64_t s = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; s < n;
       s += static_cast<int64_t>(gridDim.x) * kBlock) {
    float sum = 0.0f;
    for (int32_t e = trow[s], end = trow[s + 1]; e < end; ++e) {
      const float d = delta[tcol[e]];
      if (d > 0.0f) sum = sum + d;
    }
    if (sum > 0.0f) residual[s] = sum;
  }
}

// Jacobi form of computePRTopological (PageRank-pull.cpp:205-232).
__global__ void __launch_bounds__(kBlock) ref_topo_kernel(const int32_t* __restrict__ trow,
                                                          const int32_t* __restrict__ tcol,
                                                          const float* __restrict__ x, float* __restrict__ xn,
                                                          const uint32_t* __restrict__ nout, int64_t n, float alpha,
                                                          float base, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[kBlock / kWave];
  double diff = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * kBlock) {
    float sum = 0.0f;
    for (int32_t e = trow[v], end = trow[v + 1]; e < end; ++e) {
      const int32_t u = tcol[e];
      sum = sum + x[u] / static_cast<float>(nout[u]);
    }
    const float val = __fadd_rn(__fmul_rn(sum, alpha), base);
    diff += fabs(static_cast<double>(val) - static_cast<double>(x[v]));
    xn[v] = val;
  }
  diff = block_sum(diff, red);
  if (threadIdx.x == 0) part[blockIdx.x] = diff;
}

__global__ void sum_parts_kernel(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double red[kBlock / kWave];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *out = s;
}

}  // namespace

extern "C" gx_status gx_pagerank_pull_ref(gx_graph* g, int variant, float alpha, float tol, uint32_t max_iter,
                                          float* pr, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(pr, "pr");
  GX_REQUIRE(variant == 0 || variant == 1, "variant must be 0 (residual) or 1 (topological)");
  GX_REQUIRE(g->n > 0, "graph has no vertices");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  build_transpose(g);
  ensure_outdeg(g);
  const uint32_t* nout = g->outdeg_buf.as<uint32_t>();
  const unsigned grid = grid_for(n, kBlock, kSweepGrid);
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t it = 0, launches = 0;
  if (variant == 0) {
    float* residual = g->scratch<float>(WS_V0, n);
    float* delta = g->scratch<float>(WS_V1, n);
    unsigned* accum = g->scratch<unsigned>(WS_COUNTERS, 1);
    unsigned* host = static_cast<unsigned*>(g->host_pinned(sizeof(unsigned)));
    fill_kernel<float><<<grid, kBlock, 0, s>>>(pr, 0.0f, n);
    fill_kernel<float><<<grid, kBlock, 0, s>>>(residual, 1.0f - alpha, n);
    launches += 2;
    for (;;) {
      fill_kernel<unsigned><<<1, 64, 0, s>>>(accum, 0u, 1);
      ktimer.begin();
      ref_residual_delta_kernel<<<grid, kBlock, 0, s>>>(pr, residual, delta, nout, n, alpha, tol, accum);
      ref_residual_sum_kernel<<<grid, kBlock, 0, s>>>(g->trow(), g->tcol(), delta, residual, n);
      ktimer.end();
      GX_LAUNCHED("ref_residual_sum_kernel");
      launches += 3;
      GX_HIP(hipMemcpyAsync(host, accum, sizeof(unsigned), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      ++it;
      if (it >= max_iter || *host == 0) break;
    }
  } else {
    float* xb = g->scratch<float>(WS_V0, 2 * n);
    double* part = g->scratch<double>(WS_REDUCE, grid + 1);
    double* host = static_cast<double*>(g->host_pinned(sizeof(double)));
    float* x[2] = {xb, xb + n};
    const float base = (1.0f - alpha) / static_cast<float>(n);
    fill_kernel<float><<<grid, kBlock, 0, s>>>(x[0], 1.0f / static_cast<float>(n), n);
    launches += 1;
    int cur = 0;
    for (;;) {
      ktimer.begin();
      ref_topo_kernel<<<grid, kBlock, 0, s>>>(g->trow(), g->tcol(), x[cur], x[cur ^ 1], nout, n, alpha, base, part);
      ktimer.end();
      GX_LAUNCHED("ref_topo_kernel");
      sum_parts_kernel<<<1, kBlock, 0, s>>>(part, static_cast<int>(grid), part + grid);
      launches += 2;
      GX_HIP(hipMemcpyAsync(host, part + grid, sizeof(double), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      cur ^= 1;
      ++it;
      if (*host <= tol || it >= max_iter) break;
    }
    GX_HIP(hipMemcpyAsync(pr, x[cur], n * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = it;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(g->m) * it;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}

// ------------------------------------------------------ synchronous push
namespace {

struct PushCounters {
  unsigned count[2];
  unsigned hub_count;
  unsigned overflow;
  unsigned long long edges;
};

// Active vertices with residual above tol move it into their value and
// become pushers (delta = alpha*residual/outdeg); PageRank-push.cpp:118-150.
__global__ void __launch_bounds__(kBlock) push_take_kernel(const uint32_t* __restrict__ active, PushCounters* c,
                                                           int in, float* __restrict__ value,
                                                           float* __restrict__ residual, float* __restrict__ delta,
                                                           const int32_t* __restrict__ row, float alpha, float tol,
                                                           uint32_t* __restrict__ pushers, int out, unsigned cap) {
#pragma clang fp contract(off)
  const unsigned cnt = c->count[in];
  for (unsigned b = blockIdx.x * kBlock; b < cnt; b += gridDim.x * kBlock) {
    const unsigned i = b + threadIdx.x;
    bool push = false;
    uint32_t s = 0;
    if (i < cnt) {
      s = active[i];
      const float r = residual[s];
      if (r > tol) {
        value[s] = value[s] + r;
        residual[s] = 0.0f;
        const int32_t nout = row[s + 1] - row[s];
        if (nout > 0) {
          delta[s] = (r * alpha) / static_cast<float>(nout);
          push = true;
        }
      }
    }
    wave_append<uint32_t>(push, s, pushers, &c->count[out], cap, &c->overflow);
  }
}

struct PushOp {
  const int32_t* col;
  const float* delta;
  float* residual;
  uint32_t* stamp;
  uint32_t* next;
  PushCounters* c;
  float tol;
  uint32_t round;
  int out;
  unsigned cap;
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
#pragma clang fp contract(off)
    bool act = false;
    uint32_t d = 0;
    if (valid) {
      d = static_cast<uint32_t>(col[e]);
      const float dl = delta[u];
      const float old = atomicAdd(&residual[d], dl);
      if (old <= tol && old + dl >= tol) act = atomicExch(&stamp[d], round) != round;
    }
    wave_append<uint32_t>(act, d, next, &c->count[out], cap, &c->overflow);
  }
};

__global__ void __launch_bounds__(kBlock) push_spread_kernel(const int32_t* __restrict__ row, PushOp op,
                                                             const uint32_t* __restrict__ pushers, int in,
                                                             uint32_t* __restrict__ hubq) {
  expand_queue<kBlock>(row, pushers, op.c->count[in], 4096u, hubq, &op.c->hub_count, op.cap, &op.c->overflow, op);
}

__global__ void __launch_bounds__(kBlock) push_hubs_kernel(const int32_t* __restrict__ row, PushOp op,
                                                           const uint32_t* __restrict__ hubq) {
  expand_hubs<kBlock>(row, hubq, op.c->hub_count, op);
}

__global__ void iota_kernel(uint32_t* __restrict__ q, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    q[i] = static_cast<uint32_t>(i);
}

__global__ void push_reset_kernel(PushCounters* c, unsigned first) {
  if (threadIdx.x == 0) {
    c->count[0] = first;
    c->count[1] = 0;
    c->hub_count = 0;
    c->overflow = 0;
    c->edges = 0;
  }
}

__global__ void push_set_kernel(PushCounters* c, int idx, unsigned v) {
  if (threadIdx.x == 0) {
    c->count[idx] = v;
    c->hub_count = 0;
  }
}

}  // namespace

extern "C" gx_status gx_pagerank_push(gx_graph* g, float alpha, float tol, uint32_t max_iter, float* pr,
                                      gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(pr, "pr");
  GX_REQUIRE(g->n > 0, "graph has no vertices");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(n);
  uint32_t* active = g->scratch<uint32_t>(WS_Q0, n);
  uint32_t* pushers = g->scratch<uint32_t>(WS_Q1, n);
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, n);
  float* residual = g->scratch<float>(WS_V0, n);
  float* delta = g->scratch<float>(WS_V1, n);
  uint32_t* stamp = g->scratch<uint32_t>(WS_V2, n);
  PushCounters* c = g->scratch<PushCounters>(WS_COUNTERS, 1);
  PushCounters* host = static_cast<PushCounters*>(g->host_pinned(sizeof(PushCounters)));
  const unsigned grid = grid_for(n, kBlock, kSweepGrid);
  RegionTimer region(s);
  KernelTimer ktimer(s);
  fill_kernel<float><<<grid, kBlock, 0, s>>>(pr, 0.0f, n);
  fill_kernel<float><<<grid, kBlock, 0, s>>>(residual, 1.0f - alpha, n);
  fill_kernel<uint32_t><<<grid, kBlock, 0, s>>>(stamp, 0xffffffffu, n);
  iota_kernel<<<grid, kBlock, 0, s>>>(active, n);
  push_reset_kernel<<<1, 64, 0, s>>>(c, static_cast<unsigned>(n));
  GX_LAUNCHED("push_reset_kernel");
  uint32_t launches = 5, it = 0;
  unsigned count = static_cast<unsigned>(n);
  for (; count > 0 && it < max_iter; ++it) {
    push_set_kernel<<<1, 64, 0, s>>>(c, 1, 0u);
    ktimer.begin();
    push_take_kernel<<<std::min<unsigned>(kSweepGrid, grid_for(count, kBlock)), kBlock, 0, s>>>(
        active, c, 0, pr, residual, delta, g->row, alpha, tol, pushers, 1, cap);
    push_set_kernel<<<1, 64, 0, s>>>(c, 0, 0u);
    PushOp op{g->col, delta, residual, stamp, active, c, tol, it, 0, cap};
    push_spread_kernel<<<kSweepGrid, kBlock, 0, s>>>(g->row, op, pushers, 1, hubq);
    push_hubs_kernel<<<1024, kBlock, 0, s>>>(g->row, op, hubq);
    ktimer.end();
    GX_LAUNCHED("push_hubs_kernel");
    launches += 5;
    GX_HIP(hipMemcpyAsync(host, c, sizeof(PushCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "PageRank push worklist overflow");
    count = host->count[0];
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = it;
    st->launches = launches;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
