// PageRank: normalised pull (FP32 / FP64), the reference's own
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
// FP32 iterations run the propagation-blocked sweep of pr_blocking.hpp by
// default (knob "pr.blocking"): per-edge contributions streamed through a
// static (source chunk, destination bin) layout and exact fixed-point row
// sums, instead of a random gather per edge.  FP64 iterations and
// pr.blocking = 0 use the merge-path pull sweep below.
//
// Design of the pull sweep (gfx950, HBM/latency bound, no MFMA)
//  - The pull runs over the transpose in degree order (vertices relabelled
//    by out-degree, descending), so the most-gathered y entries share cache
//    lines; results are mapped back to the caller's ids at the end.
//  - The row sums are the merge-path tile sweep of rowsweep.hpp: tiles of a
//    fixed number of row ends + edges, columns staged with 16-byte loads,
//    every gather of a tile in flight at once, per-lane walks in LDS, and
//    coalesced per-row finalisation (x', y' = x'/outdeg, residual, dangling
//    mass).  Rows spanning tiles are completed by a second small launch.
//  - FP32 mode: compensated (Neumaier) row sums in registers and
//    TwoSum-merged wave reductions; x and y stored in FP32; the residual and
//    dangling mass are accumulated in FP64 per workgroup and reduced in a
//    fixed order (deterministic).
//  - Iterations are enqueued in batches with no host round trip: each sweep
//    reads a device `converged` flag and returns at once when it is set, so
//    surplus launches after convergence are no-ops.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <type_traits>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"
#include "pr_blocking.hpp"
#include "rowsweep.hpp"

using namespace gx;

namespace {

using rs::Acc;
using rs::block_sum;
using rs::kBlock;
using rs::kSweepGrid;

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

template <class Real>
struct PrSrc {
  const Real* __restrict__ y;
  __device__ __forceinline__ Real operator()(int64_t, int32_t c) const { return y[c]; }
  __device__ __forceinline__ Real load(int32_t c) const { return y[c]; }
  __device__ __forceinline__ Real apply(int64_t, Real v) const { return v; }
};

// FP32 update x' = base + alpha*(s_hi + s_lo) with base and alpha carried as
// unevaluated FP32 pairs (hi + lo): an error-free product (FMA) and sum
// (TwoSum) leave x' nearly correctly rounded, so its FP32 storage is the only
// rounding that matters.  Rounding alpha = 0.85 itself to FP32 (relative
// 2.8e-8) would move the fixed point by about alpha/(1-alpha) times that,
// ~2e-7 in L1 on RMAT-26.
struct F32Pair {
  float hi, lo;
};
__device__ __forceinline__ F32Pair split_f64(double v) {
  const float h = static_cast<float>(v);
  return F32Pair{h, static_cast<float>(v - static_cast<double>(h))};
}
__device__ __forceinline__ float pr_update_f32(F32Pair b, F32Pair a, float sh, float sl) {
  const float p = __fmul_rn(a.hi, sh);
  const float pe = __fmaf_rn(a.hi, sh, -p);
  float t, te;
  two_sum(p, b.hi, t, te);
  return t + (((pe + te) + b.lo) + (a.hi * sl + a.lo * sh));
}

// x'[r] = base + alpha*sum; residual |x'-x| and dangling mass in FP64.
template <class Real>
struct PrFin {
  const PrState* state;
  const Real* __restrict__ x_old;
  Real* __restrict__ x_new;
  Real* __restrict__ y_new;
  const uint32_t* __restrict__ outdeg;
  const Real* __restrict__ part;  // row sums of the hot pass (null: single pass)
  Real base, alpha;
  F32Pair bp, ap;  // FP32: base and alpha as pairs
  __device__ __forceinline__ void load() {
    base = static_cast<Real>(state->base);
    alpha = static_cast<Real>(state->alpha);
    bp = split_f64(state->base);
    ap = split_f64(state->alpha);
  }
  struct Pre {
    Real xo, pa;
    uint32_t d;
  };
  __device__ __forceinline__ Pre pre(uint32_t r) const {
    return Pre{rs::ld_stream(&x_old[r]), part ? rs::ld_stream(&part[r]) : Real(0), rs::ld_stream(&outdeg[r])};
  }
  __device__ __forceinline__ void operator()(uint32_t r, Real sum, const Pre& p, double& resid, double& dang) const {
    Real xn;
    if constexpr (std::is_same<Real, float>::value) {
      float sh, sl;
      two_sum(p.pa, sum, sh, sl);
      xn = pr_update_f32(bp, ap, sh, sl);
    } else {
      xn = base + alpha * (p.pa + sum);
    }
    resid += fabs(static_cast<double>(xn) - static_cast<double>(p.xo));
    __builtin_nontemporal_store(xn, &x_new[r]);
    const uint32_t d = p.d;
    if (d) {
      y_new[r] = xn / static_cast<Real>(d);
    } else {
      y_new[r] = Real(0);
      dang += static_cast<double>(xn);
    }
  }
};

// Hot pass of the split pull: the row sum over hot sources is stored.
template <class Real>
struct PartFin {
  Real* __restrict__ part;
  struct Pre {};
  __device__ __forceinline__ void load() {}
  __device__ __forceinline__ Pre pre(uint32_t) const { return Pre{}; }
  __device__ __forceinline__ void operator()(uint32_t r, Real sum, const Pre&, double&, double&) const {
    __builtin_nontemporal_store(sum, &part[r]);
  }
};

// Row finalisation of the propagation-blocked FP32 sweep: the row sum
// arrives exact (64-bit fixed point) as an FP32 pair; x' = base + alpha*sum
// by the compensated FP32 update; residual and dangling mass in FP64.  The
// next sweep's phase 1 derives y' = x'/outdeg itself; rows from `nonzero` on
// have no out-edges (the tail of the degree order) and carry dangling mass.
struct PbFin {
  const PrState* state;
  const float* __restrict__ x_old;
  float* __restrict__ x_new;
  int64_t nonzero;
  F32Pair base, alpha;
  __device__ __forceinline__ void load() {
    base = split_f64(state->base);
    alpha = split_f64(state->alpha);
  }
  __device__ __forceinline__ void operator()(int64_t r, unsigned long long fixed, double& resid, double& dang) const {
    float sh, sl;
    pb::fixed_to_float2(fixed, sh, sl);
    const float xn = pr_update_f32(base, alpha, sh, sl);
    resid += fabs(static_cast<double>(xn) - static_cast<double>(rs::ld_stream(&x_old[r])));
    __builtin_nontemporal_store(xn, &x_new[r]);
    if (r >= nonzero) dang += static_cast<double>(xn);
  }
};

// Row finalisation of the propagation-blocked FP64 sweep: x' = base +
// alpha*sum in FP64, as PrFin<double>; x' only, as PbFin.
struct PbFin64 {
  const PrState* state;
  const double* __restrict__ x_old;
  double* __restrict__ x_new;
  int64_t nonzero;
  double base, alpha;
  __device__ __forceinline__ void load() {
    base = state->base;
    alpha = state->alpha;
  }
  __device__ __forceinline__ void operator()(int64_t r, double sum, double& resid, double& dang) const {
    const double xn = base + alpha * sum;
    resid += fabs(xn - rs::ld_stream(&x_old[r]));
    __builtin_nontemporal_store(xn, &x_new[r]);
    if (r >= nonzero) dang += xn;
  }
};

// Hot/cold split: per row, the number of in-edges from sources < hot (a
// prefix: sources are ascending within a row) and the rest.
__global__ void split_count_kernel(const int32_t* __restrict__ rowoff, const int32_t* __restrict__ col, int64_t n,
                                   int32_t hot, uint32_t* __restrict__ ca, uint32_t* __restrict__ cb) {
  for (int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; r <= n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (r == n) {
      ca[n] = cb[n] = 0;
      continue;
    }
    const int32_t b = rowoff[r], e = rowoff[r + 1];
    int32_t lo = b, hi = e;  // first edge with col >= hot
    while (lo < hi) {
      const int32_t mid = lo + ((hi - lo) >> 1);
      if (col[mid] < hot) lo = mid + 1; else hi = mid;
    }
    ca[r] = static_cast<uint32_t>(lo - b);
    cb[r] = static_cast<uint32_t>(e - lo);
  }
}

// One wavefront per row copies its hot prefix and cold suffix.
__global__ void __launch_bounds__(kBlock) split_copy_kernel(const int32_t* __restrict__ rowoff,
                                                            const int32_t* __restrict__ col, int64_t n,
                                                            const int32_t* __restrict__ rowA,
                                                            const int32_t* __restrict__ rowB,
                                                            int32_t* __restrict__ colA, int32_t* __restrict__ colB) {
  const int64_t waves = static_cast<int64_t>(gridDim.x) * (kBlock / kWave);
  const unsigned lane = lane_id();
  for (int64_t r = blockIdx.x * static_cast<int64_t>(kBlock / kWave) + threadIdx.x / kWave; r < n; r += waves) {
    const int32_t b = rowoff[r], e = rowoff[r + 1];
    const int32_t a0 = rowA[r], k = rowA[r + 1] - a0, b0 = rowB[r];
    for (int32_t i = static_cast<int32_t>(lane); i < e - b; i += kWave) {
      const int32_t c = col[b + i];
      if (i < k)
        colA[a0 + i] = c;
      else
        colB[b0 + i - k] = c;
    }
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
  if (threadIdx.x == 0) part_dang[blockIdx.x] = dang;
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

// pr[v] = x[new_of_old[v]]: back to the caller's vertex order.
template <class Real>
__global__ void pr_unpermute_kernel(const Real* __restrict__ x, const uint32_t* __restrict__ new_of_old,
                                    Real* __restrict__ pr, int64_t n) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x)
    pr[v] = x[new_of_old[v]];
}

}  // namespace

// Hot/cold split of the degree-ordered transpose at source id `hot`.
static void build_split(gx_graph* g, int64_t hot) {
  DegreeOrder& D = g->degree_order;
  if (D.hot == hot) return;
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  DevBuf ca, cb;
  ca.ensure((n + 1) * 4);
  cb.ensure((n + 1) * 4);
  split_count_kernel<<<grid_for(n + 1, kBlock, 8192), kBlock, 0, s>>>(
      D.trow.as<int32_t>(), D.tcol.as<int32_t>(), n, static_cast<int32_t>(hot), ca.as<uint32_t>(), cb.as<uint32_t>());
  GX_LAUNCHED("split_count_kernel");
  D.rowA.ensure((n + 1) * 4);
  D.rowB.ensure((n + 1) * 4);
  exclusive_scan_u32(ca.as<uint32_t>(), D.rowA.as<uint32_t>(), n + 1, s);
  exclusive_scan_u32(cb.as<uint32_t>(), D.rowB.as<uint32_t>(), n + 1, s);
  uint32_t ma = 0;
  read_bytes(g, &ma, D.rowA.as<uint32_t>() + n, 4);
  D.colA.ensure(std::max<int64_t>(ma, 1) * 4);
  D.colB.ensure(std::max<int64_t>(m - ma, 1) * 4);
  split_copy_kernel<<<4096, kBlock, 0, s>>>(D.trow.as<int32_t>(), D.tcol.as<int32_t>(), n, D.rowA.as<int32_t>(),
                                            D.rowB.as<int32_t>(), D.colA.as<int32_t>(), D.colB.as<int32_t>());
  GX_LAUNCHED("split_copy_kernel");
  GX_HIP(hipStreamSynchronize(s));
  D.planA.ready = D.planB.ready = false;
  D.hot = hot;
  D.hot_edges = ma;
}

// Sources gathered in the hot pass (knob "pr.hot_sources"; default 0 = one
// pass, which measured faster on RMAT-26; the autotuner may pick a split).
static int64_t hot_sources(int64_t n, size_t real_bytes) {
  (void)real_bytes;
  const int64_t h = tune_effective("pr.hot_sources", n, 0);
  return (h <= 0 || h >= n) ? 0 : h;
}

// Knob "pr.blocking": 1 (default) = propagation-blocked sweep for FP32,
// 2 = also for FP64 (its sums then depend on the addition order in the last
// bit), 0 = merge-path pull for both.  Graphs outside the layout's table
// limits always use the pull.
static bool use_blocking(const gx_graph* g, size_t real_bytes) {
  const int64_t k = tune_effective("pr.blocking", g->n, g->m);
  const bool on = real_bytes == sizeof(float) ? k != 0 : k == 2;
  return on && pb::supported(g->n, g->m);
}

// Phase 1, phase 2 and the sliced-bin finalisation of one blocked sweep.
template <class Real, class Fin>
static uint32_t launch_blocked(const PbLayout& L, const Real* x, const uint32_t* outdeg,
                               typename pb::PbAcc<Real>::V* vals, int64_t n, const Fin& fin,
                               double* part_resid, double* part_dang, const unsigned* stop, hipStream_t s) {
  using Acc = typename pb::PbAcc<Real>::T;
  const pb::Job* jobs = L.jobs.as<pb::Job>();
  pb::gather_kernel<Real><<<static_cast<unsigned>(L.npieces), pb::kBlock, 0, s>>>(
      L.pieces.as<pb::Piece>(), L.gsrc16.as<uint16_t>(), x, outdeg, n, vals, stop);
  GX_LAUNCHED("pb::gather_kernel");
  pb::accum_kernel<Real, Fin><<<static_cast<unsigned>(L.ntasks), pb::kBlock, 0, s>>>(
      L.tasks.as<pb::Task>(), jobs, jobs + L.njobs[0], jobs + L.njobs[0] + L.njobs[1], vals, L.dst16.as<uint16_t>(),
      L.gacc.as<Acc>(), n, fin, part_resid, part_dang, stop);
  GX_LAUNCHED("pb::accum_kernel");
  if (L.nmulti == 0) return 2;
  pb::multi_kernel<Real, Fin><<<static_cast<unsigned>(L.nmulti), pb::kBlock, 0, s>>>(
      L.multi.as<pb::Multi>(), L.gacc.as<Acc>(), n, fin, part_resid + L.ntasks, part_dang + L.ntasks, stop);
  GX_LAUNCHED("pb::multi_kernel");
  return 3;
}

// Normalised pull PageRank over the degree-ordered transpose.
template <class Real>
static void run_pull(gx_graph* g, double alpha, double tol, uint32_t max_iter, Real* pr, gx_stats* st) {
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  build_degree_order(g);
  DegreeOrder& D = g->degree_order;
  const bool blocked = use_blocking(g, sizeof(Real));
  const int64_t hot = blocked ? 0 : hot_sources(n, sizeof(Real));
  rs::RowPlan P{}, PA{};
  if (blocked) {
    pb::build_layout(g);
    D.pb.vals.ensure(D.pb.ngroups * sizeof(Real));  // grows once for FP64 contributions
  } else if (hot > 0) {
    build_split(g, hot);
    PA = rs::build_plan(g, D.rowA.as<int32_t>(), D.colA.as<int32_t>(), n, D.hot_edges, D.planA);
    P = rs::build_plan(g, D.rowB.as<int32_t>(), D.colB.as<int32_t>(), n, g->m - D.hot_edges, D.planB);
  } else {
    P = rs::build_plan(g, D.trow.as<int32_t>(), D.tcol.as<int32_t>(), n, g->m, g->pull_plan);
  }
  const PbLayout& L = D.pb;
  Real* part = hot > 0 ? g->scratch<Real>(WS_V3, n) : nullptr;
  Real* carryA = hot > 0 ? g->scratch<Real>(WS_V4, 3 * PA.ntiles) : nullptr;
  double* partsA = hot > 0 ? g->scratch<double>(WS_V5, 2 * rs::kParts) : nullptr;
  const uint32_t* outdeg = D.outdeg.as<uint32_t>();
  Real* xb = g->scratch<Real>(WS_V0, 2 * n);
  Real* yb = g->scratch<Real>(WS_V1, 2 * n);
  Real* x[2] = {xb, xb + n};
  Real* y[2] = {yb, yb + n};
  const int nparts = blocked ? static_cast<int>(std::max<int64_t>(L.ntasks + L.nmulti, rs::kParts)) : rs::kParts;
  double* parts = g->scratch<double>(WS_REDUCE, 2 * nparts);
  double* part_resid = parts;
  double* part_dang = parts + nparts;
  Real* carry = blocked ? nullptr : g->scratch<Real>(WS_V2, 3 * P.ntiles);
  PrState* state = g->scratch<PrState>(WS_COUNTERS, 1);
  PrState* host = static_cast<PrState*>(g->host_pinned(sizeof(PrState)));

  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0;
  pr_state_init_kernel<<<1, 64, 0, s>>>(state, alpha, tol, 1.0 / static_cast<double>(n), max_iter);
  pr_init_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(x[0], y[0], outdeg, n, part_dang);
  GX_LAUNCHED("pr_init_kernel");
  pr_init_reduce_kernel<<<1, kBlock, 0, s>>>(state, part_dang, kSweepGrid);
  GX_LAUNCHED("pr_init_reduce_kernel");
  launches += 3;
  uint32_t it = 0;
  // Iterations are queued 8 at a time between host checks; sweeps queued
  // after convergence return at once and are not counted in the timings.
  const uint32_t batch = 8;
  int cur = 0;
  while (it < max_iter) {
    const uint32_t end = std::min(max_iter, it + batch);
    for (; it < end; ++it) {
      const int nxt = cur ^ 1;
      if (blocked) {
        ktimer.begin();
        if constexpr (std::is_same<Real, float>::value) {
          PbFin fin{state, x[cur], x[nxt], L.nonzero, {}, {}};
          launches += launch_blocked<float>(L, x[cur], outdeg, L.vals.as<uint32_t>(), n, fin, part_resid,
                                            part_dang, &state->converged, s);
        } else {
          PbFin64 fin{state, x[cur], x[nxt], L.nonzero, 0.0, 0.0};
          launches += launch_blocked<double>(L, x[cur], outdeg, L.vals.as<double>(), n, fin, part_resid,
                                             part_dang, &state->converged, s);
        }
        ktimer.end();
        pr_reduce_kernel<<<1, kBlock, 0, s>>>(state, part_resid, part_dang, static_cast<int>(L.ntasks + L.nmulti),
                                              static_cast<unsigned>(nxt));
        GX_LAUNCHED("pr_reduce_kernel");
        launches += 1;
        cur = nxt;
        continue;
      }
      PrSrc<Real> src{y[cur]};
      PrFin<Real> fin{state, x[cur], x[nxt], y[nxt], outdeg, part, Real(0), Real(0)};
      ktimer.begin();
      if (hot > 0) {
        rs::launch_sweep<Real, rs::HotCache<Real>::kHot>(PA, src, PartFin<Real>{part}, &state->converged, partsA,
                                                         partsA + rs::kParts, carryA, s, nullptr);
        launches += 2;
      }
      // Degree order puts the most-gathered sources first: cache them in LDS
      // (only the single pass / hot pass; the cold pass has no such prefix).
      const int np = hot > 0 ? rs::launch_sweep<Real>(P, src, fin, &state->converged, part_resid, part_dang,
                                                      carry, s, nullptr)
                             : rs::launch_sweep<Real, rs::HotCache<Real>::kHot>(
                                   P, src, fin, &state->converged, part_resid, part_dang, carry, s, nullptr);
      ktimer.end();
      pr_reduce_kernel<<<1, kBlock, 0, s>>>(state, part_resid, part_dang, np, static_cast<unsigned>(nxt));
      GX_LAUNCHED("pr_reduce_kernel");
      launches += 3;
      cur = nxt;
    }
    GX_HIP(hipMemcpyAsync(host, state, sizeof(PrState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->converged) break;
  }
  GX_HIP(hipMemcpyAsync(host, state, sizeof(PrState), hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  pr_unpermute_kernel<Real><<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(
      x[host->final_buf], D.new_of_old.as<uint32_t>(), pr, n);
  GX_LAUNCHED("pr_unpermute_kernel");
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = host->iters;
    st->launches = launches + 1;
    st->edges_visited = static_cast<uint64_t>(g->m) * host->iters;
    st->items = static_cast<uint64_t>(n) * host->iters;
    st->residual = host->resid;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms(host->iters);  // the sweeps that ran
    st->main_kernel_launches = std::min(ktimer.count(), host->iters);
    // Algorithmic bytes of one sweep.  Pull: per row y read once, x_old and
    // the out-degree read, x' and y' written (4R + 4); 4 bytes of source id
    // per edge and the row offsets (twice, and the partial sums, with the
    // hot/cold split).  Blocked: per row x and the out-degree read by phase 1,
    // x_old read and x' written by the finaliser (3R + 4); 2 bytes of
    // destination id per edge; per group a 2-byte source id, the
    // contribution written and read back; 12 bytes per phase-2 job.
    const int64_t R = sizeof(Real), m = g->m;
    const int64_t njobs = L.njobs[0] + L.njobs[1] + L.njobs[2];
    st->main_kernel_bytes = static_cast<uint64_t>(
        blocked ? 2 * m + (2 + 2 * R) * L.ngroups + (3 * R + 4) * n + 12 * njobs
                : 4 * m + (hot > 0 ? 2 : 1) * 4 * (n + 1) + (4 * R + 4) * n + (hot > 0 ? 2 * R * n : 0));
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

extern "C" gx_status gx_pagerank_layout_info(const gx_graph* g, gx_pb_layout_info* info) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(info, "info");
  std::memset(info, 0, sizeof(*info));
  const PbLayout& L = g->degree_order.pb;
  if (L.ready) {
    info->built = 1;
    info->bins = L.nbins;
    info->groups = L.ngroups;
    info->pieces = L.npieces;
    info->tasks = L.ntasks;
    info->sliced_bins = L.nmulti;
    info->jobs_long = L.njobs[0];
    info->jobs_mid = L.njobs[1];
    info->jobs_short = L.njobs[2];
    info->task_edges = L.task_entries;
  }
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

constexpr int kQCap = 2048;
// Pushes to the ids below kPushHot (the hubs on generator and degree-ordered
// graphs, whose residual words otherwise take one returning atomic per
// in-neighbour one after another) are summed per workgroup in LDS and added
// with one returning atomic per id at the end of the launch.  Residuals only
// grow, so exactly one add takes a word from at most tol to at least tol,
// whether it adds one delta or a workgroup's sum: the activation rule and
// the set of active vertices stay those of the per-edge adds.
constexpr uint32_t kPushHot = 4096;

struct PushOp {
  const int32_t* col;
  const float* delta;
  float* residual;
  uint32_t* stamp;
  uint32_t* next;
  PushCounters* c;
  LdsQueue<kQCap>* sq;
  float tol;
  uint32_t round;
  int out;
  unsigned cap;
  float* hot;  // LDS sums of the pushes to ids below kPushHot
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
#pragma clang fp contract(off)
    bool act = false;
    uint32_t d = 0;
    if (valid) {
      d = static_cast<uint32_t>(col[e]);
      const float dl = delta[u];
      if (d < kPushHot) {
        atomicAdd(&hot[d], dl);
      } else {
        const float old = atomicAdd(&residual[d], dl);
        if (old <= tol && old + dl >= tol) act = atomicExch(&stamp[d], round) != round;
      }
    }
    lq_push(*sq, act, d);
  }
  // Workgroup-uniform.  Clears the LDS sums.
  __device__ __forceinline__ void hot_clear() {
    for (uint32_t i = threadIdx.x; i < kPushHot; i += blockDim.x) hot[i] = 0.0f;
    __syncthreads();
  }
  // Workgroup-uniform.  Adds the LDS sums (ids at or above n are never pushed
  // to; every delta is positive, so a pushed id has a positive sum).
  __device__ __forceinline__ void hot_apply() {
#pragma clang fp contract(off)
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kPushHot; i += blockDim.x) {
      const float sum = hot[i];
      bool act = false;
      if (sum > 0.0f) {
        const float old = atomicAdd(&residual[i], sum);
        if (old <= tol && old + sum >= tol) act = atomicExch(&stamp[i], round) != round;
      }
      lq_push(*sq, act, i);
      flush(false);
    }
  }
  __device__ __forceinline__ void flush(bool force) {
    lq_flush(*sq, next, &c->count[out], cap, &c->overflow, force);
  }
  __device__ __forceinline__ void after_step() { flush(false); }
};

__global__ void __launch_bounds__(kBlock) push_spread_kernel(const int32_t* __restrict__ row, PushOp op,
                                                             const uint32_t* __restrict__ pushers, int in,
                                                             uint32_t* __restrict__ hubq, uint32_t hub) {
  __shared__ LdsQueue<kQCap> sq;
  __shared__ float hot[kPushHot];
  lq_init(sq);
  op.sq = &sq;
  op.hot = hot;
  op.hot_clear();
  expand_queue<kBlock>(row, pushers, op.c->count[in], hub, hubq, &op.c->hub_count, op.cap, &op.c->overflow, op);
  op.hot_apply();
  op.flush(true);
}

__global__ void __launch_bounds__(kBlock) push_hubs_kernel(const int32_t* __restrict__ row, PushOp op,
                                                           const uint32_t* __restrict__ hubq,
                                                           const uint32_t* __restrict__ hub_off) {
  if (op.c->hub_count == 0) return;
  __shared__ LdsQueue<kQCap> sq;
  __shared__ float hot[kPushHot];
  lq_init(sq);
  op.sq = &sq;
  op.hot = hot;
  op.hot_clear();
  expand_hubs<kBlock>(row, hubq, op.c->hub_count, op.cap, hub_off, op);
  op.hot_apply();
  op.flush(true);
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
  uint32_t* hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(n));
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
    PushOp op{g->col, delta, residual, stamp, active, c, nullptr, tol, it, 0, cap, nullptr};
    push_spread_kernel<<<kSweepGrid, kBlock, 0, s>>>(g->row, op, pushers, 1, hubq, hub_degree());
    launch_hub_offsets<kBlock>(g->row, hubq, &c->hub_count, cap, hub_off, s);
    push_hubs_kernel<<<kSweepGrid, kBlock, 0, s>>>(g->row, op, hubq, hub_off);
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
