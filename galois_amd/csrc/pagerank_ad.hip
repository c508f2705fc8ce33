// Differentiable PageRank: weighted forward (pull over the transpose) and the
// hand-written adjoint (transposed power iteration on the same CSR, i.e. a
// pull over out-edges), plus the edge-gradient pass.
//
// Model (no counterpart in the reference; the forward extends
// computePRTopological, lonestar/analytics/cpu/pagerank/PageRank-pull.cpp:
// 196-251, with edge weights and a teleport vector):
//   P[u,v] = w_uv / s_u, s_u = sum_v w_uv; a row with no out-edge or s_u = 0
//   is dangling and its mass is spread uniformly;
//   x = alpha*P^T x + alpha*D/n + (1-alpha)*t,  D = sum_{dangling u} x_u.
// For a loss L with dL/dx = c, the adjoint lambda = (I - alpha*A)^{-1} c with
// A = P + (1/n) d 1^T (d = dangling indicator) satisfies
//   lambda = c + alpha*(P lambda + (sum(lambda)/n) d),
//   dL/dt = (1-alpha)*lambda,
//   dL/dw_uv = alpha * x_u * (lambda_v - (P lambda)_u) / s_u.
// Both solves are Jacobi iterations on the row sweep of rowsweep.hpp (FP32
// compensated or FP64), stopping at sum|change| < tol or max_iter.  The
// backward pass needs only the converged forward vector x.
#include <algorithm>
#include <cmath>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"
#include "rowsweep.hpp"

using namespace gx;

namespace {

using rs::block_sum;
using rs::kBlock;
using rs::kSweepGrid;

struct AdState {
  double resid;
  double acc;        // dangling mass (forward) or sum(lambda) (adjoint)
  double base;       // alpha*D/n (forward) or sum(lambda)/n (adjoint)
  double alpha;
  double tol;
  double inv_n;
  unsigned converged;
  unsigned iters;
  unsigned final_buf;
  unsigned max_iter;
};

// Per-row inverse weight sums (0 marks a dangling row), FP64 accumulation.
template <class Real>
__global__ void inv_rowsum_kernel(const int32_t* __restrict__ row, const float* __restrict__ w, int64_t n,
                                  Real* __restrict__ inv_s) {
  for (int64_t u = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; u < n;
       u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s = 0;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) s += w[e];
    inv_s[u] = s > 0 ? static_cast<Real>(1.0 / s) : Real(0);
  }
}

// Weights in transpose order: wt[j] = w[tperm[j]].
template <class Real>
__global__ void permute_weights_kernel(const float* __restrict__ w, const int32_t* __restrict__ tperm, int64_t m,
                                       Real* __restrict__ wt) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < m;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x)
    wt[j] = static_cast<Real>(w[tperm[j]]);
}

// Weights in CSR order, converted to Real.
template <class Real>
__global__ void convert_weights_kernel(const float* __restrict__ w, int64_t m, Real* __restrict__ out) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < m;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[j] = static_cast<Real>(w[j]);
}

template <class Real>
struct WSrc {  // value of edge slot e = y[c] * w[e]
  const Real* __restrict__ y;
  const Real* __restrict__ w;
  __device__ __forceinline__ Real operator()(int64_t e, int32_t c) const { return y[c] * w[e]; }
};

// Forward row: x' = (1-a) t + a D/n + a*sum; y' = x' * inv_s; residual and
// dangling mass.
template <class Real>
struct FwdFin {
  const AdState* state;
  const float* __restrict__ t;
  const Real* __restrict__ inv_s;
  const Real* __restrict__ x_old;
  Real* __restrict__ x_new;
  Real* __restrict__ y_new;
  Real base, alpha;
  __device__ __forceinline__ void load() {
    base = static_cast<Real>(state->base);
    alpha = static_cast<Real>(state->alpha);
  }
  struct Pre {
    float t;
    Real is, xo;
  };
  __device__ __forceinline__ Pre pre(uint32_t r) const { return Pre{t[r], inv_s[r], x_old[r]}; }
  __device__ __forceinline__ void operator()(uint32_t r, Real sum, const Pre& p, double& resid, double& dang) const {
    const Real xn = (Real(1) - alpha) * static_cast<Real>(p.t) + base + alpha * sum;
    resid += fabs(static_cast<double>(xn) - static_cast<double>(p.xo));
    x_new[r] = xn;
    const Real is = p.is;
    y_new[r] = xn * is;
    if (is == Real(0)) dang += static_cast<double>(xn);
  }
};

// Adjoint row u: pl = (P lambda)_u (or sum(lambda)/n when dangling);
// lambda' = c + a*pl; residual and sum(lambda').  With `final_pass` set it
// only records pl (for the edge gradient) and leaves lambda unchanged.
template <class Real>
struct AdjFin {
  const AdState* state;
  const float* __restrict__ c;
  const Real* __restrict__ inv_s;
  const Real* __restrict__ lam_old;
  Real* __restrict__ lam_new;
  Real* __restrict__ pl;
  int final_pass;
  Real base, alpha;
  __device__ __forceinline__ void load() {
    base = static_cast<Real>(state->base);
    alpha = static_cast<Real>(state->alpha);
  }
  struct Pre {
    Real is, lo;
    float c;
  };
  __device__ __forceinline__ Pre pre(uint32_t r) const {
    Pre q{inv_s[r], Real(0), 0.f};
    if (!final_pass) {
      q.lo = lam_old[r];
      q.c = c[r];
    }
    return q;
  }
  __device__ __forceinline__ void operator()(uint32_t r, Real sum, const Pre& q, double& resid, double& tot) const {
    const Real is = q.is;
    const Real p = (is == Real(0)) ? base : sum * is;
    if (final_pass) {
      pl[r] = p;
      return;
    }
    const Real ln = static_cast<Real>(q.c) + alpha * p;
    resid += fabs(static_cast<double>(ln) - static_cast<double>(q.lo));
    lam_new[r] = ln;
    tot += static_cast<double>(ln);
  }
};

// mode 0 (forward): base = alpha*acc/n.  mode 1 (adjoint): base = acc/n.
__global__ void __launch_bounds__(kBlock) ad_reduce_kernel(AdState* state, const double* __restrict__ p0,
                                                           const double* __restrict__ p1, int nparts, int mode,
                                                           unsigned newest_buf) {
  if (state->converged) return;
  __shared__ double red[kBlock / kWave];
  double r = 0, a = 0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    r += p0[i];
    a += p1[i];
  }
  r = block_sum(r, red);
  a = block_sum(a, red);
  if (threadIdx.x == 0) {
    state->resid = r;
    state->acc = a;
    state->base = (mode == 0 ? state->alpha : 1.0) * a * state->inv_n;
    state->iters += 1;
    state->final_buf = newest_buf;
    if (r < state->tol || state->iters >= state->max_iter) state->converged = 1;
  }
}

__global__ void ad_state_init_kernel(AdState* s, double alpha, double tol, double inv_n, unsigned max_iter) {
  if (threadIdx.x == 0) {
    s->resid = 0;
    s->acc = 0;
    s->base = 0;
    s->alpha = alpha;
    s->tol = tol;
    s->inv_n = inv_n;
    s->converged = max_iter == 0 ? 1u : 0u;
    s->iters = 0;
    s->final_buf = 0;
    s->max_iter = max_iter;
  }
}

// Forward start: x0 = 1/n, y0 = x0*inv_s, dangling mass -> base.
template <class Real>
__global__ void __launch_bounds__(kBlock) fwd_init_kernel(Real* __restrict__ x, Real* __restrict__ y,
                                                          const Real* __restrict__ inv_s, int64_t n,
                                                          double* __restrict__ part) {
  __shared__ double red[kBlock / kWave];
  const Real x0 = Real(1) / static_cast<Real>(n);
  double dang = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * kBlock) {
    x[v] = x0;
    y[v] = x0 * inv_s[v];
    if (inv_s[v] == Real(0)) dang += static_cast<double>(x0);
  }
  dang = block_sum(dang, red);
  if (threadIdx.x == 0) part[blockIdx.x] = dang;
}

// Adjoint start: lambda0 = c; sum(lambda0) partials.
template <class Real>
__global__ void __launch_bounds__(kBlock) adj_init_kernel(Real* __restrict__ lam, const float* __restrict__ c,
                                                          int64_t n, double* __restrict__ part) {
  __shared__ double red[kBlock / kWave];
  double tot = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(kBlock) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * kBlock) {
    lam[v] = static_cast<Real>(c[v]);
    tot += static_cast<double>(c[v]);
  }
  tot = block_sum(tot, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void ad_init_reduce_kernel(AdState* state, const double* __restrict__ part, int nparts, int mode) {
  __shared__ double red[kBlock / kWave];
  double a = 0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) a += part[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) {
    state->acc = a;
    state->base = (mode == 0 ? state->alpha : 1.0) * a * state->inv_n;
  }
}

// grad_w[e] for every CSR edge e = (u,v), load balanced over rows.
template <class Real>
struct GradOp {
  const int32_t* col;
  const Real* x;
  const Real* lam;
  const Real* pl;
  const Real* inv_s;
  Real* gw;
  Real alpha;
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid) return;
    gw[e] = alpha * x[u] * (lam[col[e]] - pl[u]) * inv_s[u];
  }
};

template <class Real>
__global__ void __launch_bounds__(kBlock) grad_edges_kernel(const int32_t* __restrict__ row, GradOp<Real> op,
                                                            unsigned n, uint32_t* __restrict__ hubq,
                                                            unsigned* hub_count, unsigned* overflow, uint32_t hub) {
  expand_queue<kBlock>(row, nullptr, n, hub, hubq, hub_count, n, overflow, op);
}

template <class Real>
__global__ void __launch_bounds__(kBlock) grad_hubs_kernel(const int32_t* __restrict__ row, GradOp<Real> op,
                                                           const uint32_t* __restrict__ hubq,
                                                           const unsigned* hub_count, unsigned cap,
                                                           const uint32_t* __restrict__ hub_off) {
  expand_hubs<kBlock>(row, hubq, *hub_count, cap, hub_off, op);
}

template <class Real>
__global__ void grad_t_kernel(const Real* __restrict__ lam, int64_t n, Real scale, Real* __restrict__ gt) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x)
    gt[v] = scale * lam[v];
}

__global__ void zero_u32_kernel(unsigned* p, int k) {
  if (static_cast<int>(threadIdx.x) < k) p[threadIdx.x] = 0;
}

const float* weights_of(gx_graph* g, const float* weights) {
  if (weights) return weights;
  GX_REQUIRE(g->flags & GX_GRAPH_FWEIGHTED, "no weights given and the graph has no float weights");
  return static_cast<const float*>(g->w);
}

// Runs the iteration loop shared by forward and adjoint: `sweep(cur, nxt)`
// enqueues one sweep and returns the partial count; reduce with `mode`.
template <class SweepFn>
uint32_t iterate(gx_graph* g, AdState* state, AdState* host, double* p0, double* p1, int mode, uint32_t max_iter,
                 SweepFn sweep, uint32_t& launches) {
  hipStream_t s = g->stream;
  uint32_t it = 0;
  const uint32_t batch = profiling_enabled() ? 1 : 8;
  int cur = 0;
  while (it < max_iter) {
    const uint32_t end = std::min(max_iter, it + batch);
    for (; it < end; ++it) {
      const int nxt = cur ^ 1;
      const int np = sweep(cur, nxt);
      ad_reduce_kernel<<<1, kBlock, 0, s>>>(state, p0, p1, np, mode, static_cast<unsigned>(nxt));
      GX_LAUNCHED("ad_reduce_kernel");
      launches += 3;
      cur = nxt;
    }
    GX_HIP(hipMemcpyAsync(host, state, sizeof(AdState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->converged) break;
  }
  GX_HIP(hipMemcpyAsync(host, state, sizeof(AdState), hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  return host->iters;
}

template <class Real>
void run_forward(gx_graph* g, const float* w, const float* t, double alpha, double tol, uint32_t max_iter, Real* x,
                 gx_stats* st) {
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  build_transpose(g);
  const rs::RowPlan P = rs::build_plan(g, g->trow(), g->tcol(), n, m, g->csc_plan);
  Real* inv_s = g->scratch<Real>(WS_V4, n);
  Real* wt = g->scratch<Real>(WS_BC_DELTA, std::max<int64_t>(m, 1));
  Real* xb = g->scratch<Real>(WS_V0, 2 * n);
  Real* yb = g->scratch<Real>(WS_V1, 2 * n);
  Real* xs[2] = {xb, xb + n};
  Real* ys[2] = {yb, yb + n};
  const int nparts = rs::kParts;
  double* parts = g->scratch<double>(WS_REDUCE, 2 * nparts);
  Real* carry = g->scratch<Real>(WS_V2, 3 * P.ntiles);
  AdState* state = g->scratch<AdState>(WS_COUNTERS, 1);
  AdState* host = static_cast<AdState*>(g->host_pinned(sizeof(AdState)));
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 6;
  inv_rowsum_kernel<Real><<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(g->row, w, n, inv_s);
  if (m) permute_weights_kernel<Real><<<grid_for(m, kBlock, 8192), kBlock, 0, s>>>(w, g->tperm(), m, wt);
  ad_state_init_kernel<<<1, 64, 0, s>>>(state, alpha, tol, 1.0 / n, max_iter);
  fwd_init_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(xs[0], ys[0], inv_s, n, parts);
  ad_init_reduce_kernel<<<1, kBlock, 0, s>>>(state, parts, kSweepGrid, 0);
  GX_LAUNCHED("forward init");
  auto sweep = [&](int cur, int nxt) {
    WSrc<Real> src{ys[cur], wt};
    FwdFin<Real> fin{state, t, inv_s, xs[cur], xs[nxt], ys[nxt], Real(0), Real(0)};
    return rs::launch_sweep<Real>(P, src, fin, &state->converged, parts, parts + nparts, carry, s,
                                  &ktimer);
  };
  const uint32_t iters = iterate(g, state, host, parts, parts + nparts, 0, max_iter, sweep, launches);
  GX_HIP(hipMemcpyAsync(x, xs[host->final_buf], n * sizeof(Real), hipMemcpyDeviceToDevice, s));
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = iters;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(m) * iters;
    st->residual = host->resid;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
}

template <class Real>
void run_adjoint(gx_graph* g, const float* w, const Real* x, const float* c, double alpha, double tol,
                 uint32_t max_iter, Real* grad_w, Real* grad_t, gx_stats* st) {
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  const rs::RowPlan P = rs::build_plan(g, g->row, g->col, n, m, g->csr_plan);
  Real* inv_s = g->scratch<Real>(WS_V4, n);
  Real* wr = g->scratch<Real>(WS_BC_DELTA, std::max<int64_t>(m, 1));
  Real* lb = g->scratch<Real>(WS_V0, 2 * n);
  Real* pl = g->scratch<Real>(WS_V1, n);
  Real* lam[2] = {lb, lb + n};
  const int nparts = rs::kParts;
  double* parts = g->scratch<double>(WS_REDUCE, 2 * nparts);
  Real* carry = g->scratch<Real>(WS_V2, 3 * P.ntiles);
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, n);
  unsigned* hubc = g->scratch<unsigned>(WS_Q3, 4);
  uint32_t* hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(n));
  AdState* state = g->scratch<AdState>(WS_COUNTERS, 1);
  AdState* host = static_cast<AdState*>(g->host_pinned(sizeof(AdState)));
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 5;
  inv_rowsum_kernel<Real><<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(g->row, w, n, inv_s);
  if (m) convert_weights_kernel<Real><<<grid_for(m, kBlock, 8192), kBlock, 0, s>>>(w, m, wr);
  ad_state_init_kernel<<<1, 64, 0, s>>>(state, alpha, tol, 1.0 / n, max_iter);
  adj_init_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(lam[0], c, n, parts);
  ad_init_reduce_kernel<<<1, kBlock, 0, s>>>(state, parts, kSweepGrid, 1);
  GX_LAUNCHED("adjoint init");
  const Real* wsrc = wr;
  auto sweep = [&](int cur, int nxt) {
    WSrc<Real> src{lam[cur], wsrc};
    AdjFin<Real> fin{state, c, inv_s, lam[cur], lam[nxt], pl, 0, Real(0), Real(0)};
    return rs::launch_sweep<Real>(P, src, fin, &state->converged, parts, parts + nparts, carry, s,
                                  &ktimer);
  };
  const uint32_t iters = iterate(g, state, host, parts, parts + nparts, 1, max_iter, sweep, launches);
  const Real* lam_f = lam[host->final_buf];
  // (P lambda) at the converged lambda: one more sweep with final_pass set.
  // Its dangling value is sum(lambda)/n, which the last reduce stored in base.
  {
    WSrc<Real> src{lam_f, wsrc};
    AdjFin<Real> fin{state, c, inv_s, lam_f, nullptr, pl, 1, Real(0), Real(0)};
    rs::launch_sweep<Real>(P, src, fin, nullptr, parts, parts + nparts, carry, s, nullptr);
    launches += 2;
  }
  zero_u32_kernel<<<1, 64, 0, s>>>(hubc, 2);
  GradOp<Real> op{g->col, x, lam_f, pl, inv_s, grad_w, static_cast<Real>(alpha)};
  grad_edges_kernel<Real><<<grid_for(n, kBlock, 256 * 16), kBlock, 0, s>>>(g->row, op, static_cast<unsigned>(n), hubq,
                                                                          hubc, hubc + 1, hub_degree());
  launch_hub_offsets<kBlock>(g->row, hubq, hubc, static_cast<unsigned>(n), hub_off, s);
  grad_hubs_kernel<Real><<<kSweepGrid, kBlock, 0, s>>>(g->row, op, hubq, hubc, static_cast<unsigned>(n), hub_off);
  grad_t_kernel<Real><<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(lam_f, n, static_cast<Real>(1.0 - alpha), grad_t);
  GX_LAUNCHED("adjoint gradients");
  launches += 4;
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = iters;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(m) * (iters + 2);
    st->residual = host->resid;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
}

}  // namespace

extern "C" gx_status gx_pagerank_weighted(gx_graph* g, const float* weights, const float* teleport, double alpha,
                                          double tol, uint32_t max_iter, int precision, void* x, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(teleport, "teleport");
  check_ptr(x, "x");
  GX_REQUIRE(g->n > 0, "graph has no vertices");
  GX_REQUIRE(alpha >= 0 && alpha < 1, "alpha must be in [0,1)");
  GX_REQUIRE(precision == GX_PR_FP32 || precision == GX_PR_FP64, "precision must be GX_PR_FP32 or GX_PR_FP64");
  const float* w = weights_of(g, weights);
  if (precision == GX_PR_FP32)
    run_forward<float>(g, w, teleport, alpha, tol, max_iter, static_cast<float*>(x), st);
  else
    run_forward<double>(g, w, teleport, alpha, tol, max_iter, static_cast<double*>(x), st);
  GX_API_END
}

extern "C" gx_status gx_pagerank_adjoint(gx_graph* g, const float* weights, const void* x, const float* c,
                                         double alpha, double tol, uint32_t max_iter, int precision, void* grad_w,
                                         void* grad_t, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(x, "x");
  check_ptr(c, "c");
  check_ptr(grad_t, "grad_t");
  GX_REQUIRE(g->m == 0 || grad_w != nullptr, "grad_w is NULL");
  GX_REQUIRE(g->n > 0, "graph has no vertices");
  GX_REQUIRE(alpha >= 0 && alpha < 1, "alpha must be in [0,1)");
  GX_REQUIRE(precision == GX_PR_FP32 || precision == GX_PR_FP64, "precision must be GX_PR_FP32 or GX_PR_FP64");
  const float* w = weights_of(g, weights);
  if (precision == GX_PR_FP32)
    run_adjoint<float>(g, w, static_cast<const float*>(x), c, alpha, tol, max_iter, static_cast<float*>(grad_w),
                       static_cast<float*>(grad_t), st);
  else
    run_adjoint<double>(g, w, static_cast<const double*>(x), c, alpha, tol, max_iter, static_cast<double*>(grad_w),
                        static_cast<double*>(grad_t), st);
  GX_API_END
}
