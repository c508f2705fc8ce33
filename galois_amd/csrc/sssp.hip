// Single-source shortest paths by bucketed delta-stepping.
//
// Reference semantics: lonestar/analytics/cpu/sssp/SSSP.cpp:119-173
// (deltaStepAlgo): bucket index dist >> shift (BFS_SSSP.h:50-58, `-delta` is
// the shift), relaxation newDist = dist[src] + w with atomicMin
// (AtomicHelpers.h:60-67), push when it lowers the distance.  The GPU app
// (lonestar/analytics/gpu/sssp/sssp.cu:170-297) splits the work into a near
// list (dist within the current bucket) and a far pile.  Shortest distances
// are the unique fixpoint, so the output is bit-exact against the CPU path.
//
// Design (gfx950)
//  - near frontier: vertices whose new distance falls in the current bucket;
//    each inner round relaxes all of them with the load-balanced expansion of
//    frontier.hpp and atomicMin on the distance word; a vertex enters a
//    round's queue once (stamp word per vertex).  A vertex is expanded only
//    if its distance changed since it was last expanded (`done` word), the
//    GPU form of the reference's stale-item test (SSSP.cpp:138).
//  - far pile: vertices improved to a later bucket; each vertex is in the far
//    pile at most once (farmark word), so the pile is bounded by n.  When the
//    near frontier drains, one pass finds the smallest live far bucket and a
//    second pass moves that bucket into the near frontier and compacts the
//    rest.
#include <algorithm>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kInf = GX_DIST_INFINITY;
constexpr uint32_t kHub = 4096;

struct SsspCounters {
  unsigned near_count[2];
  unsigned far_count[2];
  unsigned hub_count;
  unsigned overflow;
  unsigned min_bucket;
  unsigned pad;
  unsigned long long edges;
  unsigned long long items;
};

__global__ void sssp_init_kernel(uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp,
                                 uint32_t* __restrict__ farmark, uint32_t* __restrict__ done, int64_t n,
                                 uint32_t src) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) {
    dist[v] = (v == src) ? 0u : kInf;
    stamp[v] = 0xffffffffu;
    farmark[v] = 0xffffffffu;
    done[v] = kInf;
  }
}

__global__ void sssp_reset_kernel(SsspCounters* c, uint32_t* nearq, uint32_t src) {
  if (threadIdx.x == 0) {
    c->near_count[0] = 1;
    c->near_count[1] = 0;
    c->far_count[0] = 0;
    c->far_count[1] = 0;
    c->hub_count = 0;
    c->overflow = 0;
    c->min_bucket = 0xffffffffu;
    c->edges = 0;
    c->items = 0;
    nearq[0] = src;
  }
}

__global__ void sssp_round_begin_kernel(SsspCounters* c, int nxt) {
  if (threadIdx.x == 0) {
    c->near_count[nxt] = 0;
    c->hub_count = 0;
  }
}

struct RelaxOp {
  const int32_t* col;
  const uint32_t* w;
  uint32_t* dist;
  uint32_t* stamp;
  uint32_t* farmark;
  uint32_t* nearq;
  uint32_t* farq;
  SsspCounters* c;
  int nxt, far_sel;
  uint32_t shift, bucket, round, phase;
  unsigned cap;

  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    bool to_near = false, to_far = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(col[e]);
      const uint32_t nd = dist[u] + w[e];
      if (nd < dist[v]) {
        const uint32_t old = atomicMin(&dist[v], nd);
        if (nd < old) {
          if ((nd >> shift) <= bucket)
            to_near = atomicExch(&stamp[v], round) != round;
          else
            to_far = atomicExch(&farmark[v], phase) != phase;
        }
      }
    }
    wave_append<uint32_t>(to_near, v, nearq, &c->near_count[nxt], cap, &c->overflow);
    wave_append<uint32_t>(to_far, v, farq, &c->far_count[far_sel], cap, &c->overflow);
  }
};

// Drop near items whose distance has not changed since their last expansion;
// mark the rest as expanded at their current distance.
__global__ void __launch_bounds__(kBlock) sssp_filter_kernel(const int32_t* __restrict__ row,
                                                             const uint32_t* __restrict__ qin,
                                                             SsspCounters* c, int cur,
                                                             const uint32_t* __restrict__ dist,
                                                             uint32_t* __restrict__ done, uint32_t* __restrict__ qout,
                                                             int out_sel, unsigned cap) {
  const unsigned cnt = c->near_count[cur];
  unsigned long long edges = 0, items = 0;
  for (unsigned b = blockIdx.x * kBlock; b < cnt; b += gridDim.x * kBlock) {
    const unsigned i = b + threadIdx.x;
    bool keep = false;
    uint32_t u = 0;
    if (i < cnt) {
      u = qin[i];
      const uint32_t d = dist[u];
      keep = done[u] != d;
      if (keep) {
        done[u] = d;
        edges += static_cast<unsigned long long>(row[u + 1] - row[u]);
        items += 1;
      }
    }
    wave_append<uint32_t>(keep, u, qout, &c->near_count[out_sel], cap, &c->overflow);
  }
  edges = wave_sum(edges);
  items = wave_sum(items);
  if (lane_id() == 0 && items) {
    atomicAdd(&c->edges, edges);
    atomicAdd(&c->items, items);
  }
}

__global__ void __launch_bounds__(kBlock) sssp_relax_kernel(const int32_t* __restrict__ row, RelaxOp op,
                                                            const uint32_t* __restrict__ qin, int cur,
                                                            uint32_t* __restrict__ hubq) {
  const unsigned qcount = op.c->near_count[cur];
  expand_queue<kBlock>(row, qin, qcount, kHub, hubq, &op.c->hub_count, op.cap, &op.c->overflow, op);
}

__global__ void __launch_bounds__(kBlock) sssp_hubs_kernel(const int32_t* __restrict__ row, RelaxOp op,
                                                           const uint32_t* __restrict__ hubq) {
  expand_hubs<kBlock>(row, hubq, op.c->hub_count, op);
}

// Smallest bucket among far items still above the current bucket.
__global__ void __launch_bounds__(kBlock) far_min_kernel(const uint32_t* __restrict__ farq, SsspCounters* c,
                                                         int sel, const uint32_t* __restrict__ dist, uint32_t shift,
                                                         uint32_t bucket) {
  const unsigned cnt = c->far_count[sel];
  uint32_t best = 0xffffffffu;
  for (unsigned i = blockIdx.x * kBlock + threadIdx.x; i < cnt; i += gridDim.x * kBlock) {
    const uint32_t b = dist[farq[i]] >> shift;
    if (b > bucket && b < best) best = b;
  }
  best = wave_min(best);
  if (lane_id() == 0 && best != 0xffffffffu) atomicMin(&c->min_bucket, best);
}

// Move far items of bucket `next` into the near queue, keep later ones
// (one copy per vertex), drop settled ones.
__global__ void __launch_bounds__(kBlock) far_split_kernel(const uint32_t* __restrict__ farq, int sel,
                                                           uint32_t* __restrict__ farq_out,
                                                           uint32_t* __restrict__ nearq, int near_sel,
                                                           SsspCounters* c, const uint32_t* __restrict__ dist,
                                                           uint32_t* __restrict__ stamp,
                                                           uint32_t* __restrict__ farmark, uint32_t shift,
                                                           uint32_t next, uint32_t round, uint32_t phase,
                                                           unsigned cap) {
  const unsigned cnt = c->far_count[sel];
  for (unsigned b0 = blockIdx.x * kBlock; b0 < cnt; b0 += gridDim.x * kBlock) {
    const unsigned i = b0 + threadIdx.x;
    bool to_near = false, keep = false;
    uint32_t v = 0;
    if (i < cnt) {
      v = farq[i];
      const uint32_t b = dist[v] >> shift;
      if (b == next)
        to_near = atomicExch(&stamp[v], round) != round;
      else if (b > next)
        keep = atomicExch(&farmark[v], phase) != phase;
    }
    wave_append<uint32_t>(to_near, v, nearq, &c->near_count[near_sel], cap, &c->overflow);
    wave_append<uint32_t>(keep, v, farq_out, &c->far_count[sel ^ 1], cap, &c->overflow);
  }
}

__global__ void far_begin_kernel(SsspCounters* c, int sel, int near_sel) {
  if (threadIdx.x == 0) {
    c->far_count[sel ^ 1] = 0;
    c->near_count[near_sel] = 0;
    c->min_bucket = 0xffffffffu;
  }
}

__global__ void far_clear_kernel(SsspCounters* c, int sel) {
  if (threadIdx.x == 0) c->far_count[sel] = 0;
}

}  // namespace

extern "C" gx_status gx_sssp_delta(gx_graph* g, uint32_t src, uint32_t shift, uint32_t* dist, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(dist, "dist");
  check_vertex(g, src, "source");
  GX_REQUIRE(g->flags & GX_GRAPH_WEIGHTED, "SSSP needs uint32 edge weights");
  GX_REQUIRE(shift < 32, "shift must be < 32");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(std::max<int64_t>(n, 1));
  uint32_t* nearq[2] = {g->scratch<uint32_t>(WS_Q0, cap), g->scratch<uint32_t>(WS_Q1, cap)};
  uint32_t* farq[2] = {g->scratch<uint32_t>(WS_Q2, cap), g->scratch<uint32_t>(WS_Q3, cap)};
  uint32_t* hubq = g->scratch<uint32_t>(WS_V0, cap);
  uint32_t* stamp = g->scratch<uint32_t>(WS_V1, cap);
  uint32_t* farmark = g->scratch<uint32_t>(WS_V2, cap);
  uint32_t* done = g->scratch<uint32_t>(WS_V3, cap);
  uint32_t* filtered = g->scratch<uint32_t>(WS_V4, cap);
  SsspCounters* c = g->scratch<SsspCounters>(WS_COUNTERS, 1);
  SsspCounters* host = static_cast<SsspCounters*>(g->host_pinned(sizeof(SsspCounters)));

  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 2;
  sssp_init_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(dist, stamp, farmark, done, n, src);
  GX_LAUNCHED("sssp_init_kernel");
  sssp_reset_kernel<<<1, 64, 0, s>>>(c, nearq[0], src);
  GX_LAUNCHED("sssp_reset_kernel");

  const unsigned relax_grid = 256 * 8;
  uint32_t bucket = 0, round = 0, phase = 0, rounds_total = 0;
  int cur = 0, far_sel = 0;
  unsigned near_count = 1, far_count = 0;
  // Each inner round: filter nearq[cur] into `filtered` (counter
  // near_count[cur^1]), then relax `filtered`, appending the next round's
  // items to nearq[cur] (counter near_count[cur]).
  for (;;) {
    while (near_count > 0) {
      // Filter stale items into `filtered` (counter near_count[cur^1]).
      const int nxt = cur ^ 1;
      sssp_round_begin_kernel<<<1, 64, 0, s>>>(c, nxt);
      sssp_filter_kernel<<<std::min<unsigned>(relax_grid, grid_for(near_count, kBlock)), kBlock, 0, s>>>(
          g->row, nearq[cur], c, cur, dist, done, filtered, nxt, cap);
      GX_LAUNCHED("sssp_filter_kernel");
      // Relax the filtered items; their count is now near_count[nxt]; the
      // relax writes the next round's items into nearq[cur] / near_count[cur].
      ++round;
      RelaxOp op{g->col, static_cast<const uint32_t*>(g->w), dist, stamp, farmark, nearq[cur], farq[far_sel], c,
                 cur, far_sel, shift, bucket, round, phase, cap};
      sssp_round_begin_kernel<<<1, 64, 0, s>>>(c, cur);
      ktimer.begin();
      sssp_relax_kernel<<<std::min<unsigned>(relax_grid, grid_for(near_count, kBlock)), kBlock, 0, s>>>(
          g->row, op, filtered, nxt, hubq);
      GX_LAUNCHED("sssp_relax_kernel");
      sssp_hubs_kernel<<<1024, kBlock, 0, s>>>(g->row, op, hubq);
      ktimer.end();
      GX_LAUNCHED("sssp_hubs_kernel");
      launches += 5;
      ++rounds_total;
      GX_HIP(hipMemcpyAsync(host, c, sizeof(SsspCounters), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      if (host->overflow) fail(GX_ERR_CAPACITY, "SSSP queue overflow");
      near_count = host->near_count[cur];
      far_count = host->far_count[far_sel];
    }
    if (far_count == 0) break;
    // Next bucket = smallest live far bucket.
    far_begin_kernel<<<1, 64, 0, s>>>(c, far_sel, cur);
    far_min_kernel<<<std::min<unsigned>(relax_grid, grid_for(far_count, kBlock)), kBlock, 0, s>>>(
        farq[far_sel], c, far_sel, dist, shift, bucket);
    GX_LAUNCHED("far_min_kernel");
    GX_HIP(hipMemcpyAsync(host, c, sizeof(SsspCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    launches += 2;
    if (host->min_bucket == 0xffffffffu) break;  // every far item is settled
    bucket = host->min_bucket;
    ++round;
    ++phase;
    far_split_kernel<<<std::min<unsigned>(relax_grid, grid_for(far_count, kBlock)), kBlock, 0, s>>>(
        farq[far_sel], far_sel, farq[far_sel ^ 1], nearq[cur], cur, c, dist, stamp, farmark, shift, bucket, round,
        phase, cap);
    GX_LAUNCHED("far_split_kernel");
    far_clear_kernel<<<1, 64, 0, s>>>(c, far_sel);
    launches += 2;
    far_sel ^= 1;
    GX_HIP(hipMemcpyAsync(host, c, sizeof(SsspCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "SSSP queue overflow");
    near_count = host->near_count[cur];
    far_count = host->far_count[far_sel];
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = rounds_total;
    st->launches = launches;
    st->edges_visited = host->edges;
    st->items = host->items;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
