// Breadth-first search: level-synchronous top-down BFS and the
// direction-optimising (push/pull) BFS.
//
// Reference semantics
//  - lonestar/analytics/cpu/bfs/bfs.cpp:245-293 (syncAlgo): level[src] = 0,
//    level[dst] = L+1 the first time dst is seen from a level-L vertex,
//    unreached = DIST_INFINITY (BFS_SSSP.h:31-32).  Levels are unique, so the
//    result is bit-exact whatever the order of discovery.
//  - lonestar/analytics/cpu/bfs/bfsDirectionOpt.cpp:205-342 (syncDOAlgo):
//    push while the scout count (edges out of the new frontier) is at most
//    edges_to_check/alpha; otherwise pull over in-edges against a frontier
//    bitmap, stopping at the first frontier neighbour, and keep pulling while
//    the frontier grows or holds more than n/beta vertices.
//
// Design (gfx950)
//  - Top-down expansion is block-local load-balanced: each 256-thread
//    workgroup takes 256 frontier vertices, exclusive-scans their degrees in
//    LDS and lets all 256 lanes walk the concatenated edge list (owner found by
//    binary search over the LDS offsets), so a wavefront never idles on a short
//    adjacency.  Vertices above the hub threshold are deferred to a hub queue
//    whose edges are spread over the whole grid.  Discovery is an atomicCAS on
//    the level word (each vertex enters the next frontier once); frontier
//    appends are wavefront-aggregated (one atomic per wavefront).
//  - Bottom-up steps test a 1-bit-per-vertex frontier bitmap (n/8 bytes, L2
//    resident for the graphs here) and stop at the first hit.
#include <algorithm>

#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kInf = GX_DIST_INFINITY;
constexpr uint32_t kNoParent = 0xffffffffu;

// Device counters (one 64-byte line).
struct BfsCounters {
  unsigned next_count;       // vertices appended to the next frontier queue
  unsigned overflow;         // a queue bound was hit
  unsigned hub_count;        // hubs deferred this level
  unsigned found;            // vertices discovered by a pull step
  unsigned long long scout;  // sum of degrees of newly discovered vertices
  unsigned pad[10];
};

__global__ void bfs_init_kernel(uint32_t* __restrict__ level, uint32_t* __restrict__ parent, int64_t n,
                                uint32_t src) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) {
    level[v] = (v == src) ? 0u : kInf;
    if (parent) parent[v] = (v == src) ? src : kNoParent;
  }
}

__global__ void reset_counters_kernel(BfsCounters* c, uint32_t* qin, uint32_t src, int seed_queue) {
  if (threadIdx.x == 0) {
    c->next_count = seed_queue ? 1u : 0u;
    c->overflow = 0;
    c->hub_count = 0;
    c->found = 0;
    c->scout = 0;
    if (seed_queue) qin[0] = src;
  }
}

__global__ void clear_level_counters_kernel(BfsCounters* c) {
  if (threadIdx.x == 0) {
    c->next_count = 0;
    c->hub_count = 0;
    c->found = 0;
    c->scout = 0;
  }
}

// Claim u for level `nl` (first writer wins); returns true for the winner.
__device__ __forceinline__ bool claim(uint32_t* level, uint32_t u, uint32_t nl) {
  if (level[u] != kInf) return false;
  return atomicCAS(&level[u], kInf, nl) == kInf;
}

__device__ __forceinline__ void discovered(bool won, uint32_t u, uint32_t from, const int32_t* __restrict__ row,
                                           uint32_t* __restrict__ parent, uint32_t* __restrict__ qout,
                                           BfsCounters* c, unsigned cap) {
  if (won && parent) parent[u] = from;
  unsigned long long d = won ? static_cast<unsigned long long>(row[u + 1] - row[u]) : 0ull;
  d = wave_sum(d);
  if (lane_id() == 0 && d) atomicAdd(&c->scout, d);
  wave_append<uint32_t>(won, u, qout, &c->next_count, cap, &c->overflow);
}

// Top-down expansion of frontier qin[0..qcount) into level nl.
__global__ void __launch_bounds__(kBlock) td_expand_kernel(
    const int32_t* __restrict__ row, const int32_t* __restrict__ col, uint32_t* __restrict__ level,
    uint32_t* __restrict__ parent, const uint32_t* __restrict__ qin, const BfsCounters* __restrict__ cin,
    uint32_t* __restrict__ qout, uint32_t* __restrict__ hubq, BfsCounters* c, uint32_t nl, uint32_t hub_thresh,
    unsigned cap) {
  __shared__ int32_t s_start[kBlock];
  __shared__ uint32_t s_off[kBlock];
  __shared__ uint32_t s_src[kBlock];
  __shared__ uint32_t s_scan[kBlock / kWave + 1];
  const unsigned qcount = cin->next_count;
  const unsigned tid = threadIdx.x;
  for (unsigned base = blockIdx.x * kBlock; base < qcount; base += gridDim.x * kBlock) {
    unsigned i = base + tid;
    uint32_t v = 0, deg = 0;
    int32_t start = 0;
    bool hub = false;
    if (i < qcount) {
      v = qin[i];
      start = row[v];
      deg = static_cast<uint32_t>(row[v + 1] - start);
      if (deg > hub_thresh) {
        hub = true;
        deg = 0;
      }
    }
    wave_append<uint32_t>(hub, v, hubq, &c->hub_count, cap, &c->overflow);
    uint32_t total;
    uint32_t excl = block_exclusive_scan<kBlock>(deg, s_scan, &total);
    s_start[tid] = start;
    s_off[tid] = excl;
    s_src[tid] = v;
    __syncthreads();
    // Block-uniform trip count so the wavefront reductions in discovered()
    // always run with every lane active.
    for (uint32_t e0 = 0; e0 < total; e0 += kBlock) {
      const uint32_t e = e0 + tid;
      bool won = false;
      uint32_t u = 0, from = 0;
      if (e < total) {
        int lo = 0, hi = kBlock - 1;
        while (lo < hi) {
          int mid = (lo + hi + 1) >> 1;
          if (s_off[mid] <= e) lo = mid; else hi = mid - 1;
        }
        u = static_cast<uint32_t>(col[s_start[lo] + (e - s_off[lo])]);
        from = s_src[lo];
        won = claim(level, u, nl);
      }
      discovered(won, u, from, row, parent, qout, c, cap);
    }
    __syncthreads();
  }
}

// Edges of deferred hub vertices, spread over the whole grid.
__global__ void __launch_bounds__(kBlock) td_hubs_kernel(const int32_t* __restrict__ row,
                                                         const int32_t* __restrict__ col,
                                                         uint32_t* __restrict__ level, uint32_t* __restrict__ parent,
                                                         const uint32_t* __restrict__ hubq,
                                                         uint32_t* __restrict__ qout, BfsCounters* c, uint32_t nl,
                                                         unsigned cap) {
  const unsigned hubs = c->hub_count;
  for (unsigned h = 0; h < hubs; ++h) {
    uint32_t v = hubq[h];
    int32_t start = row[v], end = row[v + 1];
    for (int64_t e0 = start + blockIdx.x * static_cast<int64_t>(kBlock); e0 < end;
         e0 += static_cast<int64_t>(gridDim.x) * kBlock) {
      const int64_t e = e0 + threadIdx.x;
      bool won = false;
      uint32_t u = 0;
      if (e < end) {
        u = static_cast<uint32_t>(col[e]);
        won = claim(level, u, nl);
      }
      discovered(won, u, v, row, parent, qout, c, cap);
    }
  }
}

__global__ void queue_to_bitmap_kernel(const uint32_t* __restrict__ q, const BfsCounters* __restrict__ c,
                                       uint32_t* __restrict__ bm) {
  const unsigned cnt = c->next_count;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    uint32_t v = q[i];
    atomicOr(&bm[v >> 5], 1u << (v & 31));
  }
}

// Bottom-up step: every unreached vertex looks for a frontier in-neighbour.
__global__ void __launch_bounds__(kBlock) bu_step_kernel(const int32_t* __restrict__ trow,
                                                         const int32_t* __restrict__ tcol,
                                                         const int32_t* __restrict__ row,
                                                         uint32_t* __restrict__ level, uint32_t* __restrict__ parent,
                                                         const uint32_t* __restrict__ front,
                                                         uint32_t* __restrict__ next, BfsCounters* c, int64_t n,
                                                         uint32_t nl) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v - threadIdx.x % kWave < n; v += stride) {
    bool hit = false;
    uint32_t from = 0;
    if (v < n && level[v] == kInf) {
      for (int32_t e = trow[v], end = trow[v + 1]; e < end; ++e) {
        uint32_t u = static_cast<uint32_t>(tcol[e]);
        if (front[u >> 5] & (1u << (u & 31))) {
          hit = true;
          from = u;
          break;
        }
      }
    }
    if (hit) {
      level[v] = nl;
      if (parent) parent[v] = from;
      atomicOr(&next[v >> 5], 1u << (v & 31));
    }
    unsigned long long d = hit ? static_cast<unsigned long long>(row[v + 1] - row[v]) : 0ull;
    d = wave_sum(d);
    uint64_t mask = __ballot(hit);
    if (lane_id() == 0 && mask) {
      atomicAdd(&c->found, static_cast<unsigned>(__popcll(mask)));
      atomicAdd(&c->scout, d);
    }
  }
}

__global__ void bitmap_to_queue_kernel(const uint32_t* __restrict__ bm, int64_t n, uint32_t* __restrict__ q,
                                       BfsCounters* c, unsigned cap) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v - threadIdx.x % kWave < n; v += stride) {
    bool set = v < n && (bm[v >> 5] & (1u << (v & 31)));
    wave_append<uint32_t>(set, static_cast<uint32_t>(v), q, &c->next_count, cap, &c->overflow);
  }
}

struct BfsState {
  uint32_t* q[2];
  uint32_t* hubq;
  BfsCounters* ctr[2];
  uint32_t* bm[2];
};

// Grid size for vertex-parallel sweeps: enough workgroups to fill 256 CUs
// several times over.
unsigned sweep_grid(int64_t n) { return grid_for(n, kBlock, 256 * 16); }

}  // namespace

static gx_status run_bfs(gx_graph* g, uint32_t src, bool direction_opt, uint32_t alpha, uint32_t beta,
                         uint32_t* level, uint32_t* parent, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(level, "level");
  check_vertex(g, src, "source");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  if (alpha == 0) alpha = 15;
  if (beta == 0) beta = 18;
  const int32_t* trow = g->row;
  const int32_t* tcol = g->col;
  if (direction_opt && !(g->flags & GX_GRAPH_SYMMETRIC)) {
    build_transpose(g);
    trow = g->trow();
    tcol = g->tcol();
  }
  const unsigned cap = static_cast<unsigned>(n);
  BfsState S;
  S.q[0] = g->scratch<uint32_t>(WS_Q0, n);
  S.q[1] = g->scratch<uint32_t>(WS_Q1, n);
  S.hubq = g->scratch<uint32_t>(WS_Q2, n);
  BfsCounters* ctr = g->scratch<BfsCounters>(WS_COUNTERS, 2);
  S.ctr[0] = ctr;
  S.ctr[1] = ctr + 1;
  const int64_t words = (n + 31) / 32;
  if (direction_opt) {
    S.bm[0] = g->scratch<uint32_t>(WS_BITMAP0, words);
    S.bm[1] = g->scratch<uint32_t>(WS_BITMAP1, words);
  }
  BfsCounters* host = static_cast<BfsCounters*>(g->host_pinned(sizeof(BfsCounters)));

  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0;
  bfs_init_kernel<<<sweep_grid(n), kBlock, 0, s>>>(level, parent, n, src);
  GX_LAUNCHED("bfs_init_kernel");
  // Frontier 0 = {src}: ctr[0].next_count = 1.
  reset_counters_kernel<<<1, 64, 0, s>>>(S.ctr[0], S.q[0], src, 1);
  reset_counters_kernel<<<1, 64, 0, s>>>(S.ctr[1], S.q[1], src, 0);
  launches += 3;
  int32_t rsrc[2];
  read_bytes(g, rsrc, g->row + src, sizeof(rsrc));
  int64_t scout = rsrc[1] - rsrc[0];
  int64_t edges_to_check = g->m;
  uint64_t edges_visited = scout;
  uint64_t items = 1;
  uint32_t lvl = 0;
  int cur = 0;
  unsigned frontier = 1;
  const unsigned td_grid = 256 * 8;
  while (frontier > 0) {
    if (direction_opt && scout > edges_to_check / static_cast<int64_t>(alpha)) {
      // Switch to pull: queue -> bitmap, then pull until the frontier shrinks.
      fill_kernel<uint32_t><<<sweep_grid(words), kBlock, 0, s>>>(S.bm[0], 0u, words);
      queue_to_bitmap_kernel<<<grid_for(frontier, kBlock, 4096), kBlock, 0, s>>>(S.q[cur], S.ctr[cur], S.bm[0]);
      launches += 2;
      int b = 0;
      uint64_t work = static_cast<uint64_t>(scout), old_work;
      do {
        old_work = work;
        ++lvl;
        fill_kernel<uint32_t><<<sweep_grid(words), kBlock, 0, s>>>(S.bm[b ^ 1], 0u, words);
        clear_level_counters_kernel<<<1, 64, 0, s>>>(S.ctr[0]);
        ktimer.begin();
        bu_step_kernel<<<sweep_grid(n), kBlock, 0, s>>>(trow, tcol, g->row, level, parent, S.bm[b], S.bm[b ^ 1],
                                                        S.ctr[0], n, lvl);
        ktimer.end();
        GX_LAUNCHED("bu_step_kernel");
        launches += 3;
        GX_HIP(hipMemcpyAsync(host, S.ctr[0], sizeof(BfsCounters), hipMemcpyDeviceToHost, s));
        GX_HIP(hipStreamSynchronize(s));
        work = host->found;
        items += host->found;
        edges_visited += host->scout;
        b ^= 1;
      } while (work > 0 && (work >= old_work || work > static_cast<uint64_t>(n) / beta));
      // Bitmap -> queue for the next push level.
      cur = 0;
      clear_level_counters_kernel<<<1, 64, 0, s>>>(S.ctr[0]);
      bitmap_to_queue_kernel<<<sweep_grid(n), kBlock, 0, s>>>(S.bm[b], n, S.q[0], S.ctr[0], cap);
      GX_LAUNCHED("bitmap_to_queue_kernel");
      launches += 2;
      GX_HIP(hipMemcpyAsync(host, S.ctr[0], sizeof(BfsCounters), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      if (host->overflow) fail(GX_ERR_CAPACITY, "BFS frontier queue overflow");
      frontier = host->next_count;
      scout = 1;
      continue;
    }
    ++lvl;
    edges_to_check -= scout;
    const int nxt = cur ^ 1;
    clear_level_counters_kernel<<<1, 64, 0, s>>>(S.ctr[nxt]);
    ktimer.begin();
    td_expand_kernel<<<std::min<unsigned>(td_grid, grid_for(frontier, kBlock)), kBlock, 0, s>>>(
        g->row, g->col, level, parent, S.q[cur], S.ctr[cur], S.q[nxt], S.hubq, S.ctr[nxt], lvl, 4096u, cap);
    GX_LAUNCHED("td_expand_kernel");
    td_hubs_kernel<<<1024, kBlock, 0, s>>>(g->row, g->col, level, parent, S.hubq, S.q[nxt], S.ctr[nxt], lvl, cap);
    ktimer.end();
    GX_LAUNCHED("td_hubs_kernel");
    launches += 3;
    GX_HIP(hipMemcpyAsync(host, S.ctr[nxt], sizeof(BfsCounters), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "BFS frontier queue overflow");
    frontier = host->next_count;
    items += frontier;
    scout = static_cast<int64_t>(host->scout);
    edges_visited += host->scout;
    cur = nxt;
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = lvl;
    st->launches = launches;
    st->edges_visited = edges_visited;
    st->items = items;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}

extern "C" gx_status gx_bfs(gx_graph* g, uint32_t src, uint32_t* level, gx_stats* st) {
  return run_bfs(g, src, false, 0, 0, level, nullptr, st);
}

extern "C" gx_status gx_bfs_do(gx_graph* g, uint32_t src, uint32_t alpha, uint32_t beta, uint32_t* level,
                               uint32_t* parent, gx_stats* st) {
  return run_bfs(g, src, true, alpha, beta, level, parent, st);
}
