// Multi-source shortest paths: up to 64 sources per wavefront, one lane per
// source, all groups of 64 in the same launches.
//
// Reference semantics per source: lonestar/analytics/cpu/sssp/SSSP.cpp:
// 119-173 (delta-stepping, bucket = dist >> shift, relaxation dist[src] + w
// with atomicMin, AtomicHelpers.h:60-67); the distances are the unique
// shortest-path fixpoint, so every source's row is bit-exact against the
// single-source path.  The reference runs one source per execution; batching
// sources is this library's own (no counterpart).
//
// Design (gfx950)
//  - Per group of 64 sources and vertex: 64 distances (lane j = source j, 256
//    contiguous bytes) and a 64-bit pending mask (lanes whose distance at the
//    vertex improved and has not been pushed along its out-edges yet),
//    double-buffered between rounds.
//  - A round is a sweep over the out-edge chunks of the CSR (at most
//    kEdgeChunk edges each; a hub's chunks go to many wavefronts): a
//    wavefront reads its vertex's pending mask, skips it when empty, and
//    pushes the pending lanes whose distance lies below the phase's limit
//    along the chunk's edges, eight edges in flight: per edge the 64 lanes
//    load the target's distances (one coalesced row), lower them with
//    atomicMin and set the improved lanes in the target's next pending mask
//    (one 64-bit atomicOr per edge).  Lanes at or above the limit stay
//    pending (carried by the vertex's first chunk).
//  - Phases share one bucket range across the 64 sources: the limit is
//    (first bucket + width) << shift; a phase ends when no pushed lane lands
//    below it, the next one starts at the smallest pending bucket.  Width is
//    fixed by the knob sssp.batch_width, or (0) adapts as in gx_sssp_delta
//    (knob sssp.max_merge).
//  - The round loop is driven on the device (a state word holds the buffer
//    selector, the limit and the stop flag) and captured in HIP graphs of
//    kRounds rounds, replayed until the state says done.
//  - Output is source-major: dist[i * n + v] for source i.
#include <algorithm>

#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kInf = GX_DIST_INFINITY;
constexpr int kRounds = 8;
constexpr int kUnroll = 8;
constexpr unsigned long long kMergeLow = 1ull << 14;
constexpr unsigned long long kMergeHigh = 1ull << 17;

struct BState {
  unsigned long long near;     // lanes pushed below the limit this round
  unsigned long long pushed;   // (vertex, lane) pushes admitted in the current phase
  unsigned long long edges;    // edge visits (per lane pushed)
  unsigned far_min;            // smallest pending bucket at or above the limit
  unsigned bucket;             // first bucket of the phase
  unsigned width, max_width;   // buckets of the phase; most allowed
  unsigned adapt;              // width adapts to the pushes of a phase
  unsigned limit;              // (bucket + width) << shift, saturated
  unsigned sel;                // pending masks read this round: P[sel]
  unsigned done;
  unsigned rounds;
  unsigned phases;
};

struct BBufs {
  const int32_t* row;
  const int32_t* col;
  const uint32_t* w;
  const uint32_t* cvtx;   // chunk owner
  const int32_t* cbegin;  // chunk first edge
  int64_t nchunks;
  int64_t n;
  int groups;
  uint32_t* dist;         // [group][vertex][64]
  uint64_t* pend[2];      // [group][vertex]
  BState* S;
  uint32_t shift;
};

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(x), lane));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(x >> 32), lane));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned sat_limit(unsigned bucket, unsigned width, unsigned shift) {
  const uint64_t l = (static_cast<uint64_t>(bucket) + width) << shift;
  return l > 0xffffffffull ? 0xffffffffu : static_cast<unsigned>(l);
}

__global__ void sb_init_kernel(uint32_t* __restrict__ dist, uint64_t* __restrict__ p0, uint64_t* __restrict__ p1,
                               int64_t gn) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < gn * 64;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    dist[i] = kInf;
    if (i < gn) {
      p0[i] = 0ull;
      p1[i] = 0ull;
    }
  }
}

__global__ void sb_seed_kernel(const uint32_t* __restrict__ src, int count, int64_t n, uint32_t* __restrict__ dist,
                               uint64_t* __restrict__ p0, BState* S, unsigned width, unsigned max_width,
                               unsigned shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    S->near = 0;
    S->pushed = 0;
    S->edges = 0;
    S->far_min = 0xffffffffu;
    S->bucket = 0;
    S->width = width ? width : 1;
    S->max_width = max_width;
    S->adapt = width ? 0u : 1u;
    S->limit = sat_limit(0, S->width, shift);
    S->sel = 0;
    S->done = 0;
    S->rounds = 0;
    S->phases = 1;
  }
  if (i >= count) return;
  const int64_t g = i / 64, j = i % 64;
  const uint32_t s = src[i];
  dist[(g * n + s) * 64 + j] = 0u;
  atomicOr(reinterpret_cast<unsigned long long*>(&p0[g * n + s]), 1ull << j);
}

// Clears the pending masks the coming round writes.
__global__ void sb_clear_kernel(BBufs B) {
  const BState* S = B.S;
  if (S->done) return;
  uint64_t* __restrict__ nxt = B.pend[S->sel ^ 1];
  const int64_t gn = static_cast<int64_t>(B.groups) * B.n;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < gn;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    nxt[i] = 0ull;
}

// One round over every (group, chunk).
__global__ void __launch_bounds__(kBlock) sb_relax_kernel(BBufs B) {
  BState* S = B.S;
  if (S->done) return;
  const unsigned lane = lane_id();
  const int64_t wave = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
  const unsigned sel = S->sel, limit = S->limit, shift = B.shift;
  const uint64_t* __restrict__ cur = B.pend[sel];
  uint64_t* __restrict__ nxt = B.pend[sel ^ 1];
  uint32_t* __restrict__ dist = B.dist;
  unsigned far_min = 0xffffffffu;
  unsigned long long near = 0, pushed = 0, edges = 0;
  const int64_t items = B.nchunks * B.groups;
  // 64 items per wavefront pass: every lane reads one item's pending mask,
  // then the wavefront works through the items that have one.
  for (int64_t base = wave * kWave; base < items; base += nwaves * kWave) {
    const int64_t my = base + lane;
    uint32_t mu = 0;
    uint64_t mp = 0;
    int64_t mg = 0, mc = 0;
    if (my < items) {
      mg = my / B.nchunks;
      mc = my - mg * B.nchunks;
      mu = B.cvtx[mc];
      mp = cur[mg * B.n + mu];
    }
    uint64_t live = __ballot(mp != 0);
    while (live) {
      const int i = __ffsll(static_cast<unsigned long long>(live)) - 1;
      live &= live - 1;
      const uint32_t u = __builtin_amdgcn_readlane(mu, i);
      const uint64_t pend = readlane64(mp, i);
      // chunk index < 2^32 (m < 2^31 edges), group < 64
      const int64_t c = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(mc), i));
      const int64_t g = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<uint32_t>(mg), i));
      const int64_t goff = g * B.n;
      const int32_t eb = B.cbegin[c];
      const int32_t ee = min(eb + kEdgeChunk, B.row[u + 1]);
      const uint32_t du = dist[(goff + u) * 64 + lane];
      const bool mine = (pend >> lane) & 1ull;
      const uint64_t act = __ballot(mine && du < limit);
      if (eb == B.row[u]) {  // the vertex's first chunk carries the lanes left pending
        const uint64_t keep = pend & ~act;
        if (keep) {
          if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&nxt[goff + u]), keep);
          if ((keep >> lane) & 1ull) far_min = min(far_min, du >> shift);
        }
        pushed += (act >> lane) & 1ull;
      }
      if (act == 0) continue;
      const bool on = (act >> lane) & 1ull;
      for (int32_t e0 = eb; e0 < ee; e0 += kUnroll) {
        uint32_t v[kUnroll], wt[kUnroll], dv[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
          const int32_t e = e0 + k;
          v[k] = e < ee ? static_cast<uint32_t>(B.col[e]) : 0u;
          wt[k] = e < ee ? B.w[e] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kUnroll; ++k)
          dv[k] = (on && e0 + k < ee) ? dist[(goff + v[k]) * 64 + lane] : 0u;
#pragma unroll
        for (int k = 0; k < kUnroll; ++k) {
          if (e0 + k < ee) {  // wavefront-uniform
            const uint32_t nd = du + wt[k];
            bool imp = false;
            if (on && nd < dv[k]) imp = nd < atomicMin(&dist[(goff + v[k]) * 64 + lane], nd);
            const uint64_t im = __ballot(imp);
            if (im) {
              if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&nxt[goff + v[k]]), im);
              if (imp) {
                if (nd < limit) near += 1;
                else far_min = min(far_min, nd >> shift);
              }
            }
          }
        }
        edges += on ? static_cast<unsigned long long>(min(kUnroll, ee - e0)) : 0ull;
      }
    }
  }
  // Reduced over the workgroup first: one atomic per state word per
  // workgroup rather than per wavefront.
  constexpr int kW = kBlock / kWave;
  __shared__ unsigned s_min[kW];
  __shared__ unsigned long long s_sum[3][kW];
  far_min = wave_min(far_min);
  near = wave_sum(near);
  pushed = wave_sum(pushed);
  edges = wave_sum(edges);
  const unsigned w = threadIdx.x / kWave;
  if (lane == 0) {
    s_min[w] = far_min;
    s_sum[0][w] = near;
    s_sum[1][w] = pushed;
    s_sum[2][w] = edges;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned fm = 0xffffffffu;
    unsigned long long nr = 0, pu = 0, ed = 0;
    for (int i = 0; i < kW; ++i) {
      fm = min(fm, s_min[i]);
      nr += s_sum[0][i];
      pu += s_sum[1][i];
      ed += s_sum[2][i];
    }
    if (fm != 0xffffffffu) atomicMin(&S->far_min, fm);
    if (nr) atomicAdd(&S->near, nr);
    if (pu) atomicAdd(&S->pushed, pu);
    if (ed) atomicAdd(&S->edges, ed);
  }
}

// End of a round: swap the pending masks; stay in the phase while lanes were
// pushed below the limit, else start the next phase at the smallest pending
// bucket (adapting the width to the pushes of the phase), or stop.
__global__ void sb_step_kernel(BState* S, unsigned shift) {
  if (threadIdx.x != 0 || S->done) return;
  S->rounds += 1;
  S->sel ^= 1;
  if (S->near > 0) {
    S->near = 0;
    S->far_min = 0xffffffffu;
    return;
  }
  if (S->far_min == 0xffffffffu) {
    S->done = 1;
    return;
  }
  if (S->adapt) {
    if (S->pushed < kMergeLow && S->width < S->max_width) S->width = min(2 * S->width, S->max_width);
    else if (S->pushed > kMergeHigh && S->width > 1) S->width >>= 1;
  }
  S->pushed = 0;
  S->bucket = S->far_min;
  S->limit = sat_limit(S->bucket, S->width, shift);
  S->far_min = 0xffffffffu;
  S->phases += 1;
}

// dist[g][v][lane] -> out[(g * 64 + lane) * n + v] for live lanes, through an
// LDS tile of 64 vertices x 64 lanes.
__global__ void __launch_bounds__(kBlock) sb_out_kernel(const uint32_t* __restrict__ dist, int64_t n, int groups,
                                                        int count, uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[64][65];
  const int64_t tiles = (n + 63) / 64;
  for (int64_t t = blockIdx.x; t < tiles * groups; t += gridDim.x) {
    const int64_t g = t / tiles, v0 = (t - g * tiles) * 64;
    for (int k = threadIdx.x; k < 64 * 64; k += kBlock) {
      const int i = k / 64, j = k % 64;  // vertex v0 + i, lane j
      tile[i][j] = v0 + i < n ? dist[((g * n) + v0 + i) * 64 + j] : kInf;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 64 * 64; k += kBlock) {
      const int j = k / 64, i = k % 64;  // lane j, vertex v0 + i
      const int64_t srci = g * 64 + j;
      if (srci < count && v0 + i < n) out[srci * n + v0 + i] = tile[i][j];
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" gx_status gx_sssp_batched(gx_graph* g, const uint32_t* sources, uint32_t count, uint32_t shift,
                                     uint32_t* dist, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  GX_REQUIRE(count == 0 || sources != nullptr, "sources is NULL");
  GX_REQUIRE(count == 0 || dist != nullptr, "dist is NULL");
  GX_REQUIRE(g->flags & GX_GRAPH_WEIGHTED, "SSSP needs uint32 edge weights");
  GX_REQUIRE(shift < 32, "shift must be < 32");
  for (uint32_t i = 0; i < count; ++i) check_vertex(g, sources[i], "source");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0, rounds = 0;
  unsigned long long edges = 0;
  if (count > 0 && n > 0) {
    const EdgeChunks C = out_edge_chunks(g);
    // Groups per pass bounded by memory: 256 B of distances and two 8-byte
    // masks per vertex per group.
    const int64_t per_group = n * (64 * 4 + 16);
    const int64_t budget = int64_t(64) << 30;
    const int max_groups = static_cast<int>(std::min<int64_t>(64, std::max<int64_t>(1, budget / per_group)));
    const unsigned max_width = static_cast<unsigned>(tune_effective("sssp.max_merge", n, g->m));
    const unsigned width = static_cast<unsigned>(tune_effective("sssp.batch_width", n, g->m));
    uint32_t* dsrc = g->scratch<uint32_t>(WS_Q3, count);
    GX_HIP(hipMemcpyAsync(dsrc, sources, count * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    BState* S = g->scratch<BState>(WS_COUNTERS, 1);
    BState* host = static_cast<BState*>(g->host_pinned(sizeof(BState)));
    const unsigned wgrid = 256 * 16;
    for (uint32_t first = 0; first < count; first += 64u * max_groups) {
      const uint32_t cnt = std::min<uint32_t>(count - first, 64u * max_groups);
      BBufs B;
      B.row = g->row;
      B.col = g->col;
      B.w = static_cast<const uint32_t*>(g->w);
      B.cvtx = C.vtx;
      B.cbegin = C.begin;
      B.nchunks = C.count;
      B.n = n;
      B.groups = static_cast<int>((cnt + 63) / 64);
      const int64_t gn = static_cast<int64_t>(B.groups) * n;
      B.dist = g->scratch<uint32_t>(WS_BC_SIGMA, gn * 64);
      B.pend[0] = g->scratch<uint64_t>(WS_BC_MASK0, gn);
      B.pend[1] = g->scratch<uint64_t>(WS_BC_MASK1, gn);
      B.S = S;
      B.shift = shift;
      sb_init_kernel<<<grid_for(gn * 64, kBlock, 16384), kBlock, 0, s>>>(B.dist, B.pend[0], B.pend[1], gn);
      GX_LAUNCHED("sb_init_kernel");
      sb_seed_kernel<<<grid_for(cnt, kBlock), kBlock, 0, s>>>(dsrc + first, static_cast<int>(cnt), n, B.dist,
                                                               B.pend[0], S, width, max_width, shift);
      GX_LAUNCHED("sb_seed_kernel");
      launches += 2;
      uint64_t key = 0;
      for (const void* p : {(const void*)B.row, (const void*)B.col, (const void*)B.w, (const void*)B.cvtx,
                            (const void*)B.cbegin, (const void*)B.dist, (const void*)B.pend[0],
                            (const void*)B.pend[1], (const void*)S})
        key = mix_key(key, reinterpret_cast<uintptr_t>(p));
      key = mix_key(mix_key(mix_key(key, static_cast<uint64_t>(gn)), static_cast<uint64_t>(B.nchunks)), shift);
      ktimer.begin();
      for (;;) {
        g->replay(g->sssp_batched_batch, key, [&](hipStream_t cs) {
          for (int r = 0; r < kRounds; ++r) {
            sb_clear_kernel<<<grid_for(gn, kBlock, 8192), kBlock, 0, cs>>>(B);
            sb_relax_kernel<<<wgrid, kBlock, 0, cs>>>(B);
            sb_step_kernel<<<1, 64, 0, cs>>>(S, shift);
          }
          GX_LAUNCHED("sssp batched rounds");
        });
        launches += 3 * kRounds;
        GX_HIP(hipMemcpyAsync(host, S, sizeof(BState), hipMemcpyDeviceToHost, s));
        GX_HIP(hipStreamSynchronize(s));
        if (host->done) break;
      }
      ktimer.end();
      rounds += host->rounds;
      edges += host->edges;
      sb_out_kernel<<<grid_for(((n + 63) / 64) * B.groups, 1, 8192), kBlock, 0, s>>>(
          B.dist, n, B.groups, static_cast<int>(cnt), dist + static_cast<int64_t>(first) * n);
      GX_LAUNCHED("sb_out_kernel");
      ++launches;
    }
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = rounds;
    st->launches = launches;
    st->edges_visited = edges;
    st->items = count;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
