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
//  - Top-down expansion uses the load-balanced workgroup expansion of
//    frontier.hpp (degree above frontier.hub_degree, 128 by default, goes
//    to a grid-wide hub pass).  Discovery is
//    an atomicCAS on the level word, so each vertex enters the next frontier
//    once.  New vertices are staged in an LDS queue and moved to the global
//    frontier with one atomic per workgroup flush; per-lane counters are
//    reduced once per workgroup.
//  - gx_bfs drives its levels on the device: each level is a fixed kernel
//    sequence (begin, expand, hubs, end) that reads the frontier size from a
//    device state word; levels are captured in HIP graphs (7 for the first
//    replay, 8 for the next; the end of a level and the begin of the next in
//    one launch) and replayed until the frontier is empty.
//  - Bottom-up steps test a 1-bit-per-vertex frontier bitmap (n/8 bytes, L2
//    resident for the graphs here) and stop at the first hit.
#include <algorithm>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr int kQCap = 2048;
constexpr uint32_t kInf = GX_DIST_INFINITY;
constexpr uint32_t kNoParent = 0xffffffffu;
constexpr int kGrid = 256 * 8;
// Graphs without hubs (maximum degree at or below the hub threshold, such as
// road-like grids) have small frontiers per step and many steps; a quarter
// of the workgroups per launch costs less there (2048-grid SSSP 26.4 ->
// 21.7 ms, BFS 2.03 -> 1.57 ms), while RMAT frontiers need the full grid.
constexpr int kGridNoHubs = 512;
constexpr int kLevelsPerGraph = 8;
constexpr int kLevelsFirst = 7;  // first replay: power-law graphs need 5-6 levels plus the empty one

struct BfsState {
  unsigned count[2];      // frontier sizes (double buffered by q)
  unsigned hub_count;
  unsigned overflow;
  unsigned q;             // frontier read this level
  unsigned level;         // level being written
  unsigned active;        // this level has a frontier
  unsigned done;
  unsigned found;         // vertices found by a pull step
  unsigned pad;
  unsigned long long scout;  // degree sum of the vertices discovered this level
  unsigned long long edges;  // degree sum of every discovered vertex
  unsigned long long items;
  // Output arrays of the current call, read by the captured level kernels
  // (so the captured graph does not depend on the caller's buffers).
  uint32_t* level_out;
  uint32_t* parent_out;
};

struct Bufs {
  const int32_t* row;
  const int32_t* col;
  uint32_t* level;
  uint32_t* parent;
  uint32_t* q[2];
  uint32_t* hubq;
  BfsState* S;
  unsigned cap;
  uint32_t* hub_off;
  uint32_t hub;   // frontier.hub_degree
  uint32_t* vis;  // visited bitmap (top-down BFS only; null for the DO-BFS)
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

__global__ void bfs_reset_kernel(BfsState* S, uint32_t* q0, uint32_t src, uint32_t* level, uint32_t* parent,
                                 uint32_t* vis) {
  if (threadIdx.x == 0) {
    if (vis) vis[src >> 5] |= 1u << (src & 31);
    S->level_out = level;
    S->parent_out = parent;
    S->count[0] = 1;
    S->count[1] = 0;
    S->hub_count = 0;
    S->overflow = 0;
    S->q = 0;
    S->level = 0;
    S->active = 0;
    S->done = 0;
    S->found = 0;
    S->scout = 0;
    S->edges = 0;
    S->items = 1;
    q0[0] = src;
  }
}

__device__ __forceinline__ void level_begin(BfsState* S) {
  if (S->done || S->overflow || S->count[S->q] == 0) {
    S->active = 0;
    S->done = 1;
    return;
  }
  S->active = 1;
  S->level += 1;
  S->count[S->q ^ 1] = 0;
  S->hub_count = 0;
  S->scout = 0;
}

__device__ __forceinline__ void level_end(BfsState* S) {
  if (!S->active) return;
  S->q ^= 1;
  S->items += S->count[S->q];
  S->edges += S->scout;
  S->active = 0;
}

__global__ void level_begin_kernel(BfsState* S) {
  if (threadIdx.x == 0) level_begin(S);
}

__global__ void level_end_kernel(BfsState* S) {
  if (threadIdx.x == 0) level_end(S);
}

// End of one level and begin of the next in one launch (inside a captured
// batch of levels).
__global__ void level_turn_kernel(BfsState* S) {
  if (threadIdx.x != 0) return;
  level_end(S);
  level_begin(S);
}

// Plain pointers only: the output arrays come from the device state and the
// next queue is picked by the state's index, so the kernel-argument Bufs is
// never copied or modified (a modified copy, with its runtime-indexed queue
// array, would be kept in scratch memory).
struct TdOp {
  const int32_t* __restrict__ row;
  const int32_t* __restrict__ col;
  uint32_t* level;
  uint32_t* parent;
  uint32_t* qout;
  unsigned* qcount;
  unsigned cap;
  BfsState* S;
  LdsQueue<kQCap>* sq;
  uint32_t nl;
  unsigned long long scout;  // this lane's degree sum of discovered vertices
  uint32_t* vis;
  // With a visited bitmap (n/8 bytes, mostly cache-resident) an edge to an
  // already discovered vertex is dropped without the random read of its
  // level word; a bit set late only sends the edge on to the level check.
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    bool won = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(col[e]);
      const bool seen = vis && ((vis[v >> 5] >> (v & 31)) & 1u);
      if (!seen && level[v] == kInf && atomicCAS(&level[v], kInf, nl) == kInf) {
        won = true;
        if (vis) atomicOr(&vis[v >> 5], 1u << (v & 31));
        if (parent) parent[v] = u;
        scout += static_cast<unsigned long long>(row[v + 1] - row[v]);
      }
    }
    lq_push(*sq, won, v);
  }
  __device__ __forceinline__ void after_step() { lq_flush(*sq, qout, qcount, cap, &S->overflow, false); }
};

__device__ __forceinline__ TdOp make_td(const Bufs& B, LdsQueue<kQCap>* sq) {
  BfsState* S = B.S;
  const unsigned nq = S->q ^ 1;
  return TdOp{B.row, B.col, S->level_out, S->parent_out, B.q[nq], &S->count[nq], B.cap, S, sq, S->level, 0ull, B.vis};
}

// Workgroup-uniform: the scout count is summed over the workgroup, so the
// state word takes one atomic per workgroup rather than one per wavefront.
__device__ __forceinline__ void td_finish(TdOp& op) {
  lq_flush(*op.sq, op.qout, op.qcount, op.cap, &op.S->overflow, true);
  __shared__ unsigned long long s_sc[kBlock / kWave];
  const unsigned long long sc = wave_sum(op.scout);
  if (lane_id() == 0) s_sc[threadIdx.x / kWave] = sc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / kWave; ++i) t += s_sc[i];
    if (t) atomicAdd(&op.S->scout, t);
  }
}

__global__ void __launch_bounds__(kBlock) td_expand_kernel(Bufs B) {
  BfsState* S = B.S;
  if (!S->active) return;
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  TdOp op = make_td(B, &sq);
  expand_queue<kBlock>(B.row, B.q[S->q], S->count[S->q], B.hub, B.hubq, &S->hub_count, B.cap, &S->overflow, op);
  td_finish(op);
}

__global__ void __launch_bounds__(kBlock) td_hubs_kernel(Bufs B) {
  BfsState* S = B.S;
  if (!S->active || S->hub_count == 0) return;
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  TdOp op = make_td(B, &sq);
  expand_hubs<kBlock>(B.row, B.hubq, S->hub_count, B.cap, B.hub_off, op);
  td_finish(op);
}

// Expansion of one level.  `hubs` false: no vertex exceeds the hub
// threshold, so the hub list stays empty and its two launches are left out.
void enqueue_expand(const Bufs& B, bool hubs, hipStream_t s) {
  td_expand_kernel<<<hubs ? kGrid : kGridNoHubs, kBlock, 0, s>>>(B);
  if (hubs) {
    launch_hub_offsets<kBlock>(B.row, B.hubq, &B.S->hub_count, B.cap, B.hub_off, s);
    td_hubs_kernel<<<kGrid, kBlock, 0, s>>>(B);
  }
}

// One level with its begin and end (a level launched on its own).
void enqueue_level(const Bufs& B, bool hubs, hipStream_t s) {
  level_begin_kernel<<<1, 64, 0, s>>>(B.S);
  enqueue_expand(B, hubs, s);
  level_end_kernel<<<1, 64, 0, s>>>(B.S);
  GX_LAUNCHED("bfs level");
}

// `levels` consecutive levels with the end of each and the begin of the
// next in one launch.
void enqueue_levels(const Bufs& B, bool hubs, int levels, hipStream_t s) {
  level_begin_kernel<<<1, 64, 0, s>>>(B.S);
  for (int i = 0; i < levels; ++i) {
    enqueue_expand(B, hubs, s);
    if (i + 1 < levels)
      level_turn_kernel<<<1, 64, 0, s>>>(B.S);
    else
      level_end_kernel<<<1, 64, 0, s>>>(B.S);
  }
  GX_LAUNCHED("bfs levels");
}

// ------------------------------------------------------------ bottom-up
__global__ void queue_to_bitmap_kernel(const uint32_t* __restrict__ q, const unsigned* __restrict__ count,
                                       uint32_t* __restrict__ bm) {
  const unsigned cnt = *count;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t v = q[i];
    atomicOr(&bm[v >> 5], 1u << (v & 31));
  }
}

// Every unreached vertex looks for a frontier in-neighbour.
__global__ void __launch_bounds__(kBlock) bu_step_kernel(const int32_t* __restrict__ trow,
                                                         const int32_t* __restrict__ tcol,
                                                         const int32_t* __restrict__ row,
                                                         uint32_t* __restrict__ level, uint32_t* __restrict__ parent,
                                                         const uint32_t* __restrict__ front,
                                                         uint32_t* __restrict__ next, BfsState* S, int64_t n,
                                                         uint32_t nl) {
  __shared__ unsigned long long red[kBlock / kWave];
  __shared__ unsigned redc[kBlock / kWave];
  unsigned long long deg = 0;
  unsigned found = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (level[v] != kInf) continue;
    for (int32_t e = trow[v], end = trow[v + 1]; e < end; ++e) {
      const uint32_t u = static_cast<uint32_t>(tcol[e]);
      if (front[u >> 5] & (1u << (u & 31))) {
        level[v] = nl;
        if (parent) parent[v] = u;
        atomicOr(&next[v >> 5], 1u << (v & 31));
        deg += static_cast<unsigned long long>(row[v + 1] - row[v]);
        ++found;
        break;
      }
    }
  }
  deg = wave_sum(deg);
  found = wave_sum(found);
  if (lane_id() == 0) {
    red[threadIdx.x / kWave] = deg;
    redc[threadIdx.x / kWave] = found;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long d = 0;
    unsigned f = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
      d += red[i];
      f += redc[i];
    }
    if (f) {
      atomicAdd(&S->found, f);
      atomicAdd(&S->scout, d);
    }
  }
}

__global__ void __launch_bounds__(kBlock) bitmap_to_queue_kernel(const uint32_t* __restrict__ bm, int64_t n,
                                                                 uint32_t* __restrict__ q, unsigned* count,
                                                                 unsigned cap, unsigned* overflow) {
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x); b < n;
       b += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t v = b + threadIdx.x;
    const bool set = v < n && (bm[v >> 5] & (1u << (v & 31)));
    lq_push(sq, set, static_cast<uint32_t>(v));
    lq_flush(sq, q, count, cap, overflow, false);
  }
  lq_flush(sq, q, count, cap, overflow, true);
}

__global__ void do_pull_begin_kernel(BfsState* S) {
  if (threadIdx.x == 0) {
    S->found = 0;
    S->scout = 0;
  }
}

__global__ void set_level_kernel(BfsState* S, uint32_t lvl) {
  if (threadIdx.x == 0) S->level = lvl;
}

__global__ void do_to_queue_begin_kernel(BfsState* S) {
  if (threadIdx.x == 0) {
    S->q = 0;
    S->count[0] = 0;
  }
}

unsigned sweep_grid(int64_t n) { return grid_for(n, kBlock, 256 * 16); }

}  // namespace

extern "C" gx_status gx_bfs(gx_graph* g, uint32_t src, uint32_t* level, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(level, "level");
  check_vertex(g, src, "source");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(n);
  // Visited bitmap from 2^25 vertices on: below that the 4-byte level array
  // (<= 128 MB) stays cache-resident and the bitmap's clear and atomics cost
  // more than they save (RMAT-18 0.455 -> 0.52 ms, RMAT-22 2.03 -> 2.21 ms
  // with it; RMAT-26 33.9 -> 24.7 ms).
  const int64_t words = (n + 31) / 32;
  const bool use_vis = tune_effective("bfs.visited_bitmap", n, g->m) != 0;
  uint32_t* vis = use_vis ? g->scratch<uint32_t>(WS_BITMAP0, words) : nullptr;
  Bufs B{g->row, g->col, level, nullptr, {g->scratch<uint32_t>(WS_Q0, n), g->scratch<uint32_t>(WS_Q1, n)},
         g->scratch<uint32_t>(WS_Q2, n), g->scratch<BfsState>(WS_COUNTERS, 1), cap,
         g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(n)), hub_degree(), vis};
  const bool hubs = max_degree(g) > static_cast<int64_t>(B.hub);
  // Graphs with hubs run on the degree-ordered copy of the CSR
  // (frontier.relabel): the levels of the most often reached vertices then
  // share a few cached megabytes; the source is mapped in, the levels back.
  const bool relabel = hubs && tune_effective("frontier.relabel", n, g->m) != 0;
  uint32_t run_src = src;
  if (relabel) {
    build_relabeled(g);
    B.row = g->relab.row.as<int32_t>();
    B.col = g->relab.col.as<int32_t>();
    B.level = g->scratch<uint32_t>(WS_V5, cap);
    run_src = read_scalar(g, g->relab.new_of_old.as<uint32_t>() + src);
  }
  BfsState* host = static_cast<BfsState*>(g->host_pinned(sizeof(BfsState)));
  RegionTimer region(s);
  KernelTimer ktimer(s);
  bfs_init_kernel<<<sweep_grid(n), kBlock, 0, s>>>(B.level, nullptr, n, run_src);
  if (vis) fill_kernel<uint32_t><<<sweep_grid(words), kBlock, 0, s>>>(vis, 0u, words);
  bfs_reset_kernel<<<1, 64, 0, s>>>(B.S, B.q[0], run_src, B.level, nullptr, B.vis);
  GX_LAUNCHED("bfs init");
  uint64_t key = 0;
  for (const void* p : {(const void*)B.row, (const void*)B.col, (const void*)B.q[0],
                        (const void*)B.q[1], (const void*)B.hubq, (const void*)B.S, (const void*)B.hub_off,
                        (const void*)B.vis})
    key = mix_key(key, reinterpret_cast<uintptr_t>(p));
  key = mix_key(mix_key(mix_key(key, cap), B.hub), hubs ? 1u : 0u);
  uint32_t replays = 0, slots = 0;
  for (;;) {
    const int nl = replays == 0 ? kLevelsFirst : kLevelsPerGraph;
    ktimer.begin();
    g->replay(replays == 0 ? g->bfs_head : g->bfs_batch, key, [&](hipStream_t cs) {
      enqueue_levels(B, hubs, nl, cs);
    });
    ktimer.end();
    ++replays;
    slots += static_cast<uint32_t>(nl);
    GX_HIP(hipMemcpyAsync(host, B.S, sizeof(BfsState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "BFS frontier queue overflow");
    if (host->done) break;
  }
  if (relabel) unpermute_u32(B.level, g->relab.new_of_old.as<uint32_t>(), level, n, s);
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = host->level - 1;  // levels with a non-empty frontier beyond the source
    st->launches = (vis ? 3 : 2) + slots * (hubs ? 4 : 2) + replays + (relabel ? 1 : 0);
    st->edges_visited = host->edges + static_cast<uint64_t>(0);
    st->items = host->items;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  // Count the source's own adjacency in the visited edges.
  if (st) {
    int32_t r[2];
    read_bytes(g, r, g->row + src, sizeof(r));
    st->edges_visited += static_cast<uint64_t>(r[1] - r[0]);
  }
  GX_API_END
}

// Parents back to the caller's vertex ids: out[v] = old_of_new[p] for the
// parent p of v's degree-ordered id (the source's own id, or none).
__global__ void unpermute_parent_kernel(const uint32_t* __restrict__ pin, const uint32_t* __restrict__ new_of_old,
                                        const uint32_t* __restrict__ old_of_new, uint32_t* __restrict__ out,
                                        int64_t n) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint32_t p = pin[new_of_old[v]];
    out[v] = p == kNoParent ? p : old_of_new[p];
  }
}

extern "C" gx_status gx_bfs_do(gx_graph* g, uint32_t src, uint32_t alpha, uint32_t beta, uint32_t* level,
                               uint32_t* parent, gx_stats* st) {
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
  if (!(g->flags & GX_GRAPH_SYMMETRIC)) {
    build_transpose(g);
    trow = g->trow();
    tcol = g->tcol();
  }
  const unsigned cap = static_cast<unsigned>(n);
  Bufs B{g->row, g->col, level, parent, {g->scratch<uint32_t>(WS_Q0, n), g->scratch<uint32_t>(WS_Q1, n)},
         g->scratch<uint32_t>(WS_Q2, n), g->scratch<BfsState>(WS_COUNTERS, 1), cap,
         g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(n)), hub_degree(), nullptr};
  const int64_t words = (n + 31) / 32;
  uint32_t* bm[2] = {g->scratch<uint32_t>(WS_BITMAP0, words), g->scratch<uint32_t>(WS_BITMAP1, words)};
  const bool hubs = max_degree(g) > static_cast<int64_t>(B.hub);
  // Symmetric graphs with hubs run on the degree-ordered copy of the CSR
  // (frontier.relabel): a bottom-up step scans each unvisited vertex's
  // neighbours hubs first, and the hubs are the first vertices reached, so
  // most scans stop at their first neighbour; levels and parents are mapped
  // back at the end.
  const bool relabel = (g->flags & GX_GRAPH_SYMMETRIC) && hubs && tune_effective("frontier.relabel", n, g->m) != 0;
  const int32_t* row = g->row;
  uint32_t run_src = src;
  if (relabel) {
    build_relabeled(g);
    row = g->relab.row.as<int32_t>();
    trow = row;
    tcol = g->relab.col.as<int32_t>();
    B.row = row;
    B.col = tcol;
    B.level = g->scratch<uint32_t>(WS_V5, cap);
    B.parent = parent ? g->scratch<uint32_t>(WS_V4, cap) : nullptr;
    run_src = read_scalar(g, g->relab.new_of_old.as<uint32_t>() + src);
  }
  BfsState* host = static_cast<BfsState*>(g->host_pinned(sizeof(BfsState)));
  auto read_state = [&]() {
    GX_HIP(hipMemcpyAsync(host, B.S, sizeof(BfsState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "BFS frontier queue overflow");
  };
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 2;
  bfs_init_kernel<<<sweep_grid(n), kBlock, 0, s>>>(B.level, B.parent, n, run_src);
  bfs_reset_kernel<<<1, 64, 0, s>>>(B.S, B.q[0], run_src, B.level, B.parent, nullptr);
  GX_LAUNCHED("bfs init");
  int32_t rsrc[2];
  read_bytes(g, rsrc, row + run_src, sizeof(rsrc));
  int64_t scout = rsrc[1] - rsrc[0];
  int64_t edges_to_check = g->m;
  uint64_t edges_visited = static_cast<uint64_t>(scout), items = 1;
  uint32_t lvl = 0;
  read_state();
  unsigned frontier = 1;
  while (frontier > 0) {
    if (scout > edges_to_check / static_cast<int64_t>(alpha)) {
      // Pull phase: queue -> bitmap; pull while the frontier grows or stays
      // above n/beta.
      const unsigned q = host->q;
      fill_kernel<uint32_t><<<sweep_grid(words), kBlock, 0, s>>>(bm[0], 0u, words);
      queue_to_bitmap_kernel<<<grid_for(frontier, kBlock, 4096), kBlock, 0, s>>>(B.q[q], &B.S->count[q], bm[0]);
      launches += 2;
      int b = 0;
      uint64_t work = static_cast<uint64_t>(scout), old_work;
      do {
        old_work = work;
        ++lvl;
        fill_kernel<uint32_t><<<sweep_grid(words), kBlock, 0, s>>>(bm[b ^ 1], 0u, words);
        do_pull_begin_kernel<<<1, 64, 0, s>>>(B.S);
        ktimer.begin();
        bu_step_kernel<<<sweep_grid(n), kBlock, 0, s>>>(trow, tcol, row, B.level, B.parent, bm[b], bm[b ^ 1], B.S,
                                                        n, lvl);
        ktimer.end();
        GX_LAUNCHED("bu_step_kernel");
        launches += 3;
        read_state();
        work = host->found;
        items += host->found;
        edges_visited += host->scout;
        b ^= 1;
      } while (work > 0 && (work >= old_work || work > static_cast<uint64_t>(n) / beta));
      // Bitmap -> queue for the next push level.
      do_to_queue_begin_kernel<<<1, 64, 0, s>>>(B.S);
      bitmap_to_queue_kernel<<<sweep_grid(n), kBlock, 0, s>>>(bm[b], n, B.q[0], &B.S->count[0], cap,
                                                              &B.S->overflow);
      GX_LAUNCHED("bitmap_to_queue_kernel");
      launches += 2;
      read_state();
      frontier = host->count[0];
      scout = 1;
      // The push kernels take the level number from the state.
      set_level_kernel<<<1, 64, 0, s>>>(B.S, lvl);
      GX_LAUNCHED("set_level_kernel");
      continue;
    }
    edges_to_check -= scout;
    ktimer.begin();
    enqueue_level(B, hubs, s);
    ktimer.end();
    launches += hubs ? 5 : 3;
    read_state();
    ++lvl;
    frontier = host->count[host->q];
    items += frontier;
    scout = static_cast<int64_t>(host->scout);
    edges_visited += host->scout;
  }
  if (relabel) {
    unpermute_u32(B.level, g->relab.new_of_old.as<uint32_t>(), level, n, s);
    if (parent) {
      unpermute_parent_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(
          B.parent, g->relab.new_of_old.as<uint32_t>(), g->relab.old_of_new.as<uint32_t>(), parent, n);
      GX_LAUNCHED("unpermute_parent_kernel");
    }
    launches += parent ? 2 : 1;
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
