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
// appends the neighbour to the next frontier through the workgroup's LDS
// queue.  On the degree-ordered copy the decrements of the hub ids are
// counted in LDS first (kHot below).
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
  unsigned in;      // dead frontier read this round (q[in]); the next is q[in ^ 1]
  unsigned active;  // this round has a frontier
  unsigned done;
  unsigned rounds;
  unsigned long long items;
  unsigned long long alive;
  unsigned long long edges;
};

__global__ void kcore_reset_kernel(KcoreCounters* c) {
  if (threadIdx.x == 0) {
    c->count[0] = c->count[1] = 0;
    c->hub_count = 0;
    c->overflow = 0;
    c->in = 0;
    c->active = 0;
    c->done = 0;
    c->rounds = 0;
    c->items = 0;
    c->alive = 0;
    c->edges = 0;
  }
}

constexpr int kQCap = 2048;
// On the degree-ordered copy the hubs are the smallest ids (and on graphs
// whose generator gives small ids the high degrees, such as RMAT without a
// relabelling, too), and in the first round (every vertex below k dies at
// once) most decrements land on them: a hub's degree word then takes
// millions of returning atomics one after another.  Decrements of ids below
// kHot are counted per workgroup in LDS instead and applied with one
// atomicSub per counter at the end of the launch; the atomic that takes a
// word from k or above to below k finds the death, so every vertex still
// dies exactly once.
constexpr uint32_t kHot = 4096;
constexpr int kRoundsPerGraph = 4;
constexpr int kRoundsFirst = 2;  // first replay: many graphs need one round (plus the empty one)

// degree[v] = out-degree; vertices below k form the first dead frontier
// (staged in LDS, one global append per workgroup flush).  Vertices without
// edges are dead from the start but decrement no one, so they are left out
// of it (on RMAT graphs they are a large share of the vertices).
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
      dead = d < k && d > 0;
    }
    lq_push(sq, dead, static_cast<uint32_t>(v));
    lq_flush(sq, q, &c->count[0], cap, &c->overflow, false);
  }
  lq_flush(sq, q, &c->count[0], cap, &c->overflow, true);
}

// Plain pointers, the output queue picked from the device state in the
// kernel (the kernel arguments are never modified or copied).
struct PeelOp {
  const int32_t* col;
  uint32_t* degree;
  uint32_t* qout;
  KcoreCounters* c;
  LdsQueue<kQCap>* sq;
  uint32_t k;
  unsigned out;
  unsigned cap;
  uint32_t hub;   // frontier.hub_degree
  uint32_t* hot;  // LDS counters of the ids below nhot (nhot = 0: none)
  uint32_t nhot;
  __device__ __forceinline__ void operator()(bool valid, uint32_t, int32_t e) {
    bool dies = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(col[e]);
      if (v < nhot)
        atomicAdd(&hot[v], 1u);
      else
        dies = atomicSub(&degree[v], 1u) == k;
    }
    lq_push(*sq, dies, v);
  }
  __device__ __forceinline__ void flush(bool force) {
    lq_flush(*sq, qout, &c->count[out], cap, &c->overflow, force);
  }
  __device__ __forceinline__ void after_step() { flush(false); }
};

struct KcoreBufs {
  const int32_t* row;
  const int32_t* col;
  uint32_t* degree;
  uint32_t* q0;
  uint32_t* q1;
  uint32_t* hubq;
  uint32_t* hub_off;
  KcoreCounters* c;
  uint32_t k;
  unsigned cap;
  uint32_t hub;
};

__device__ __forceinline__ PeelOp make_peel(const KcoreBufs& B, LdsQueue<kQCap>* sq, uint32_t* hot, uint32_t nhot) {
  const unsigned out = B.c->in ^ 1u;
  return PeelOp{B.col, B.degree, out ? B.q1 : B.q0, B.c, sq, B.k, out, B.cap, B.hub, hot, nhot};
}

// Workgroup-uniform.  Clears the LDS counters of the hot ids.
__device__ __forceinline__ void hot_clear(uint32_t* hot) {
  for (uint32_t i = threadIdx.x; i < kHot; i += kBlock) hot[i] = 0;
  __syncthreads();
}

// Workgroup-uniform.  Applies the workgroup's counted decrements of the hot
// ids (ids at or above n are never counted).
__device__ __forceinline__ void hot_apply(PeelOp& op) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kHot; i += kBlock) {
    const uint32_t cnt = op.hot[i];
    bool dies = false;
    if (cnt) {
      const uint32_t old = atomicSub(&op.degree[i], cnt);
      dies = old >= op.k && old - cnt < op.k;
    }
    lq_push(*op.sq, dies, i);
    op.flush(false);
  }
}

template <bool kAgg>
__global__ void __launch_bounds__(kBlock) kcore_peel_kernel(KcoreBufs B) {
  KcoreCounters* c = B.c;
  if (!c->active) return;
  __shared__ LdsQueue<kQCap> sq;
  __shared__ uint32_t hot[kAgg ? kHot : 1];
  lq_init(sq);
  if constexpr (kAgg) hot_clear(hot);
  PeelOp op = make_peel(B, &sq, hot, kAgg ? kHot : 0u);
  const unsigned in = c->in;
  expand_queue<kBlock>(B.row, in ? B.q1 : B.q0, c->count[in], B.hub, B.hubq, &c->hub_count, B.cap, &c->overflow,
                       op);
  if constexpr (kAgg) hot_apply(op);
  op.flush(true);
}

template <bool kAgg>
__global__ void __launch_bounds__(kBlock) kcore_hubs_kernel(KcoreBufs B) {
  KcoreCounters* c = B.c;
  if (!c->active || c->hub_count == 0) return;
  __shared__ LdsQueue<kQCap> sq;
  __shared__ uint32_t hot[kAgg ? kHot : 1];
  lq_init(sq);
  if constexpr (kAgg) hot_clear(hot);
  PeelOp op = make_peel(B, &sq, hot, kAgg ? kHot : 0u);
  expand_hubs<kBlock>(B.row, B.hubq, c->hub_count, B.cap, B.hub_off, op);
  if constexpr (kAgg) hot_apply(op);
  op.flush(true);
}

// Round bracket, driven by the device state: a round runs while the dead
// frontier is nonempty; the end swaps the frontiers.
__device__ __forceinline__ void round_begin(KcoreCounters* c) {
  if (c->done || c->overflow || c->count[c->in] == 0) {
    c->active = 0;
    c->done = 1;
    return;
  }
  c->active = 1;
  c->count[c->in ^ 1] = 0;
  c->hub_count = 0;
  c->items += c->count[c->in];
  c->rounds += 1;
}

__device__ __forceinline__ void round_end(KcoreCounters* c) {
  if (!c->active) return;
  c->in ^= 1;
  c->active = 0;
}

__global__ void kcore_round_begin_kernel(KcoreCounters* c) {
  if (threadIdx.x == 0) round_begin(c);
}

__global__ void kcore_round_end_kernel(KcoreCounters* c) {
  if (threadIdx.x == 0) round_end(c);
}

// End of one round and begin of the next in one launch.
__global__ void kcore_round_turn_kernel(KcoreCounters* c) {
  if (threadIdx.x != 0) return;
  round_end(c);
  round_begin(c);
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
  uint32_t launches = 2;
  // Rounds are driven on the device: rounds (begin, peel, hub layout, hubs,
  // end) captured in HIP graphs, kRoundsFirst for the first replay and
  // kRoundsPerGraph for the next ones, replayed until the state says done,
  // so the host synchronises once per replay.
  KcoreBufs B{g->row, g->col, degree, q[0], q[1], hubq, hub_off, c, k, cap, hub_degree()};
  // No vertex above the hub threshold: the hub list stays empty and its two
  // launches are left out of each round.
  const bool hubs = max_degree(g) > static_cast<int64_t>(B.hub);
  // Graphs with hubs peel on the degree-ordered copy of the CSR
  // (frontier.relabel): the decrements mostly land on the hubs, whose
  // degree words then share a few cached megabytes; the degrees are mapped
  // back before the membership pass.
  const bool relabel = hubs && tune_effective("frontier.relabel", n, g->m) != 0;
  if (relabel) {
    build_relabeled(g);
    B.row = g->relab.row.as<int32_t>();
    B.col = g->relab.col.as<int32_t>();
    B.degree = g->scratch<uint32_t>(WS_V5, cap);
  }
  // LDS counting of the ids below kHot: on the degree-ordered copy, or when
  // those ids hold at least 8 times their share of the adjacency entries.
  bool agg = relabel;
  if (!agg && hubs && n > static_cast<int64_t>(kHot)) {
    int32_t low = 0;
    GX_HIP(hipMemcpyAsync(&low, B.row + kHot, sizeof(low), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    agg = static_cast<double>(low) * static_cast<double>(n) >=
          8.0 * static_cast<double>(kHot) * static_cast<double>(g->m);
  }
  kcore_reset_kernel<<<1, 64, 0, s>>>(c);
  kcore_init_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(B.row, B.degree, n, k, q[0], c, cap);
  GX_LAUNCHED("kcore_init_kernel");
  uint64_t key = 0;
  for (const void* p : {(const void*)B.row, (const void*)B.col, (const void*)B.degree, (const void*)B.q0,
                        (const void*)B.q1, (const void*)B.hubq, (const void*)B.hub_off, (const void*)B.c})
    key = mix_key(key, reinterpret_cast<uintptr_t>(p));
  key = mix_key(mix_key(mix_key(mix_key(mix_key(key, k), cap), B.hub), hubs ? 1u : 0u), agg ? 1u : 0u);
  uint32_t replays = 0, slots = 0;
  for (;;) {
    const int nr = replays == 0 ? kRoundsFirst : kRoundsPerGraph;
    ktimer.begin();
    g->replay(replays == 0 ? g->kcore_head : g->kcore_batch, key, [&](hipStream_t cs) {
      kcore_round_begin_kernel<<<1, 64, 0, cs>>>(c);
      for (int i = 0; i < nr; ++i) {
        if (agg)
          kcore_peel_kernel<true><<<256 * 8, kBlock, 0, cs>>>(B);
        else
          kcore_peel_kernel<false><<<256 * 8, kBlock, 0, cs>>>(B);
        if (hubs) {
          launch_hub_offsets<kBlock>(B.row, B.hubq, &c->hub_count, cap, B.hub_off, cs);
          if (agg)
            kcore_hubs_kernel<true><<<256 * 8, kBlock, 0, cs>>>(B);
          else
            kcore_hubs_kernel<false><<<256 * 8, kBlock, 0, cs>>>(B);
        }
        if (i + 1 < nr)
          kcore_round_turn_kernel<<<1, 64, 0, cs>>>(c);
        else
          kcore_round_end_kernel<<<1, 64, 0, cs>>>(c);
      }
      GX_LAUNCHED("kcore rounds");
    });
    ktimer.end();
    ++replays;
    slots += static_cast<uint32_t>(nr);
    GX_HIP(hipMemcpyAsync(host, c, sizeof(KcoreCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "k-core worklist overflow");
    if (host->done) break;
  }
  launches += slots * (hubs ? 4 : 2) + replays;
  const uint32_t rounds = host->rounds;
  const uint64_t items = host->items;
  uint64_t edges = 0;
  if (relabel) {
    unpermute_u32(B.degree, g->relab.new_of_old.as<uint32_t>(), degree, n, s);
    launches += 1;
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
