// k-core decomposition (symmetric graphs) by synchronous cascade peeling.
//
// Reference semantics: lonestar/analytics/cpu/k-core/kcore.cpp:86-95
// (degreeCounting: currentDegree = degree), :103-115 (initial dead worklist:
// currentDegree < k), :124-159 (syncCascadeKCore: every dead vertex
// decrements each neighbour with atomicSubtract; a neighbour whose old value
// was exactly k dies and joins the next worklist), :201-217 (alive:
// currentDegree >= k).  The D-IrGL GPU form is KCoreStep1/2
// (lonestar/analytics/distributed/k-core/kcore_push_cuda.py:10-87).  Every
// dead vertex is processed exactly once and decrements all its neighbours, so
// the final degree array equals the number of alive neighbours: bit-exact
// whatever the schedule.
//
// Design (gfx950): the dead frontier is expanded with the load-balanced
// workgroup expansion of frontier.hpp (hubs spread over the grid); each
// decrement is one atomicSub on the degree word, and the `old == k` winner
// appends the neighbour to the next frontier with a wavefront-aggregated
// append.
#include <algorithm>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;

struct KcoreCounters {
  unsigned count[2];
  unsigned hub_count;
  unsigned overflow;
  unsigned long long alive;
  unsigned long long edges;
};

__global__ void kcore_reset_kernel(KcoreCounters* c) {
  if (threadIdx.x == 0) {
    c->count[0] = c->count[1] = 0;
    c->hub_count = 0;
    c->overflow = 0;
    c->alive = 0;
    c->edges = 0;
  }
}

constexpr int kQCap = 2048;

// degree[v] = out-degree; vertices below k form the first dead frontier
// (staged in LDS, one global append per workgroup flush).
__global__ void __launch_bounds__(kBlock) kcore_init_kernel(const int32_t* __restrict__ row,
                                                            uint32_t* __restrict__ degree, int64_t n, uint32_t k,
                                                            uint32_t* __restrict__ q, KcoreCounters* c,
                                                            unsigned cap) {
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x); b < n; b += stride) {
    const int64_t v = b + threadIdx.x;
    bool dead = false;
    if (v < n) {
      const uint32_t d = static_cast<uint32_t>(row[v + 1] - row[v]);
      degree[v] = d;
      dead = d < k;
    }
    lq_push(sq, dead, static_cast<uint32_t>(v));
    lq_flush(sq, q, &c->count[0], cap, &c->overflow, false);
  }
  lq_flush(sq, q, &c->count[0], cap, &c->overflow, true);
}

struct PeelOp {
  const int32_t* col;
  uint32_t* degree;
  uint32_t* qout;
  KcoreCounters* c;
  LdsQueue<kQCap>* sq;
  uint32_t k;
  int out;
  unsigned cap;
  uint32_t hub;  // frontier.hub_degree
  __device__ __forceinline__ void operator()(bool valid, uint32_t, int32_t e) {
    bool dies = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(col[e]);
      dies = atomicSub(&degree[v], 1u) == k;
    }
    lq_push(*sq, dies, v);
  }
  __device__ __forceinline__ void flush(bool force) {
    lq_flush(*sq, qout, &c->count[out], cap, &c->overflow, force);
  }
  __device__ __forceinline__ void after_step() { flush(false); }
};

__global__ void __launch_bounds__(kBlock) kcore_peel_kernel(const int32_t* __restrict__ row, PeelOp op,
                                                            const uint32_t* __restrict__ qin, int in,
                                                            uint32_t* __restrict__ hubq) {
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  op.sq = &sq;
  const unsigned cnt = op.c->count[in];
  expand_queue<kBlock>(row, qin, cnt, op.hub, hubq, &op.c->hub_count, op.cap, &op.c->overflow, op);
  op.flush(true);
}

__global__ void __launch_bounds__(kBlock) kcore_hubs_kernel(const int32_t* __restrict__ row, PeelOp op,
                                                            const uint32_t* __restrict__ hubq,
                                                            const uint32_t* __restrict__ hub_off) {
  if (op.c->hub_count == 0) return;
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  op.sq = &sq;
  expand_hubs<kBlock>(row, hubq, op.c->hub_count, op.cap, hub_off, op);
  op.flush(true);
}

__global__ void kcore_round_begin_kernel(KcoreCounters* c, int out) {
  if (threadIdx.x == 0) {
    c->count[out] = 0;
    c->hub_count = 0;
  }
}

__global__ void __launch_bounds__(kBlock) kcore_finish_kernel(const uint32_t* __restrict__ degree,
                                                              uint8_t* __restrict__ alive, int64_t n, uint32_t k,
                                                              KcoreCounters* c) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long cnt = 0;
  for (; v < n; v += stride) {
    const bool a = degree[v] >= k;
    if (alive) alive[v] = a ? 1 : 0;
    cnt += a ? 1ull : 0ull;
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(&c->alive, cnt);
}

}  // namespace

extern "C" gx_status gx_kcore(gx_graph* g, uint32_t k, uint32_t* degree, uint8_t* alive, uint64_t* alive_count,
                              gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(degree, "degree");
  GX_REQUIRE(g->flags & GX_GRAPH_SYMMETRIC, "k-core needs a symmetric graph");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(std::max<int64_t>(n, 1));
  uint32_t* q[2] = {g->scratch<uint32_t>(WS_Q0, cap), g->scratch<uint32_t>(WS_Q1, cap)};
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, cap);
  uint32_t* hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(cap));
  KcoreCounters* c = g->scratch<KcoreCounters>(WS_COUNTERS, 1);
  KcoreCounters* host = static_cast<KcoreCounters*>(g->host_pinned(sizeof(KcoreCounters)));
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 2, rounds = 0;
  uint64_t edges = 0, items = 0;
  kcore_reset_kernel<<<1, 64, 0, s>>>(c);
  kcore_init_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(g->row, degree, n, k, q[0], c, cap);
  GX_LAUNCHED("kcore_init_kernel");
  GX_HIP(hipMemcpyAsync(host, c, sizeof(KcoreCounters), hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  unsigned count = host->count[0];
  int in = 0;
  while (count > 0) {
    items += count;
    const int out = in ^ 1;
    kcore_round_begin_kernel<<<1, 64, 0, s>>>(c, out);
    PeelOp op{g->col, degree, q[out], c, nullptr, k, out, cap, hub_degree()};
    ktimer.begin();
    kcore_peel_kernel<<<std::min<unsigned>(256 * 8, grid_for(count, kBlock)), kBlock, 0, s>>>(g->row, op, q[in], in,
                                                                                              hubq);
    GX_LAUNCHED("kcore_peel_kernel");
    launch_hub_offsets<kBlock>(g->row, hubq, &c->hub_count, cap, hub_off, s);
    kcore_hubs_kernel<<<256 * 8, kBlock, 0, s>>>(g->row, op, hubq, hub_off);
    ktimer.end();
    GX_LAUNCHED("kcore_hubs_kernel");
    launches += 3;
    ++rounds;
    GX_HIP(hipMemcpyAsync(host, c, sizeof(KcoreCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "k-core worklist overflow");
    count = host->count[out];
    in = out;
  }
  kcore_finish_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(degree, alive, n, k, c);
  GX_LAUNCHED("kcore_finish_kernel");
  launches += 1;
  GX_HIP(hipMemcpyAsync(host, c, sizeof(KcoreCounters), hipMemcpyDeviceToHost, s));
  double ms = region.stop();
  GX_HIP(hipStreamSynchronize(s));
  if (alive_count) *alive_count = host->alive;
  (void)edges;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = rounds;
    st->launches = launches;
    st->items = items;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
