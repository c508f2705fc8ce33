// Brandes betweenness centrality: one source per pass (gx_bc) and all
// sources batched, 64 per wavefront (gx_bc_batched).
//
// Reference semantics: Level BC, lonestar/analytics/cpu/
// betweennesscentrality/LevelStructs.h:87-185 (forward: BFS levels, path
// counts sigma accumulated from level-L predecessors; backward, deepest level
// first: delta[v] = sigma[v] * sum_{w succ, level L+1} (1 + delta[w]) /
// sigma[w], then bc[v] += delta[v]; the source itself is excluded), the FP64
// Outer form OuterStructs.h:41-102 and the D-IrGL kernels
// lonestar/analytics/distributed/betweennesscentrality/bc_level_cuda.py:
// 43-143.  Directed convention, not halved for symmetric graphs.  Path counts
// are FP32, dependencies FP64, bc FP64.
//
// Plain (gx_bc): per source a level-synchronous push forward (atomicCAS on
// the level word, FP32 atomicAdd of sigma) and a pull backward over each
// level's vertices (FP64 atomicAdd into the dependency of the vertex), both
// with the load-balanced expansion of frontier.hpp.  The level lists are
// kept back to back in one array, the stack of worklists of the reference.
//
// Batched (gx_bc_batched): sources are processed 64 at a time per wavefront,
// lane j owning source j of its group; several groups run in the same
// launches.  Per vertex and group the frontier of each level is one 64-bit
// mask.  Forward level L -> L+1 is a pull: a wavefront takes a chunk of a
// vertex's in-edges, loads 64 neighbour masks at once, and for every
// neighbour with a useful bit lane j adds sigma[u][j]; new bits become the
// vertex's level-(L+1) mask.  Backward is the matching pull over out-edges
// against the level-(L+1) masks.  Adjacencies longer than one chunk are
// split over several wavefronts that combine with atomics; all other
// vertices are written by a single wavefront.  A vertex needs only the lanes
// whose source lies in its own connected component (on symmetric graphs the
// component labels are known): once those lanes have reached it, it is
// skipped, so in a union of graphs a group's sweep stops touching a graph
// when its few sources there are done.  The forward level loop is driven on
// the device (a state word holds the level and the stop flag; level buffers
// come from a device table the host extends ahead of each replay) and
// captured in HIP graphs of kFwdLevels levels, so the host synchronises once
// per replay; the backward level loop has a fixed length once the forward
// sweep is done and is captured and replayed the same way.
#include <algorithm>
#include <vector>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kInf = GX_BC_INFINITY;

// ================================================================== plain
struct PlainCounters {
  unsigned count;
  unsigned hub_count;
  unsigned overflow;
  unsigned pad;
};

__global__ void bc_plain_init_kernel(uint32_t* __restrict__ level, float* __restrict__ sigma,
                                     double* __restrict__ delta, int64_t n) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    level[v] = kInf;
    sigma[v] = 0.f;
    delta[v] = 0.0;
  }
}

__global__ void bc_plain_source_kernel(uint32_t* level, float* sigma, uint32_t* order, PlainCounters* c,
                                       uint32_t s) {
  if (threadIdx.x == 0) {
    level[s] = 0;
    sigma[s] = 1.f;
    order[0] = s;
    c->count = 0;
    c->hub_count = 0;
    c->overflow = 0;
  }
}

__global__ void bc_plain_clear_kernel(PlainCounters* c) {
  if (threadIdx.x == 0) {
    c->count = 0;
    c->hub_count = 0;
  }
}

constexpr int kQCap = 2048;

// Forward discovery from level L: claim unreached neighbours for L+1 and stage
// them in LDS (the next level's list).
struct FwdOp {
  const int32_t* col;
  uint32_t* level;
  uint32_t* next;
  PlainCounters* c;
  LdsQueue<kQCap>* sq;
  uint32_t nl;
  unsigned cap;
  __device__ __forceinline__ void operator()(bool valid, uint32_t, int32_t e) {
    bool won = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(col[e]);
      won = level[v] == kInf && atomicCAS(&level[v], kInf, nl) == kInf;
    }
    lq_push(*sq, won, v);
  }
  __device__ __forceinline__ void flush(bool force) { lq_flush(*sq, next, &c->count, cap, &c->overflow, force); }
  __device__ __forceinline__ void after_step() { flush(false); }
};

// Path counts of level L+1 by a pull over in-edges: sigma[v] = sum of
// sigma[u] over in-neighbours at level L.  Consecutive lanes hold edges of the
// same vertex, so a wavefront segmented sum precedes one atomicAdd per run.
struct SigmaOp {
  const int32_t* col;
  const uint32_t* level;
  float* sigma;
  uint32_t pred;
  __device__ __forceinline__ void operator()(bool valid, uint32_t v, int32_t e) {
    float x = 0.f;
    if (valid) {
      const uint32_t u = static_cast<uint32_t>(col[e]);
      if (level[u] == pred) x = sigma[u];
    }
    bool tail;
    const float sum = wave_segmented_sum(valid ? v : 0xffffffffu, x, tail);
    if (valid && tail && sum != 0.f) atomicAdd(&sigma[v], sum);
  }
  __device__ __forceinline__ void after_step() {}
};

// Dependencies of level L: delta[v] += sum over successors w at L+1 of
// (1 + delta[w]) / sigma[w], segmented per vertex before the atomicAdd.
struct BwdOp {
  const int32_t* col;
  const uint32_t* level;
  const float* sigma;
  double* delta;
  uint32_t succ;
  __device__ __forceinline__ void operator()(bool valid, uint32_t v, int32_t e) {
    double x = 0.0;
    if (valid) {
      const uint32_t w = static_cast<uint32_t>(col[e]);
      if (level[w] == succ) x = (1.0 + delta[w]) / static_cast<double>(sigma[w]);
    }
    bool tail;
    const double sum = wave_segmented_sum(valid ? v : 0xffffffffu, x, tail);
    if (valid && tail && sum != 0.0) atomicAdd(&delta[v], sum);
  }
  __device__ __forceinline__ void after_step() {}
};

template <class Op>
__global__ void __launch_bounds__(kBlock) bc_expand_kernel(const int32_t* __restrict__ row, Op op,
                                                           const uint32_t* __restrict__ q, unsigned qcount,
                                                           uint32_t* __restrict__ hubq, PlainCounters* c,
                                                           unsigned cap, uint32_t hub) {
  expand_queue<kBlock>(row, q, qcount, hub, hubq, &c->hub_count, cap, &c->overflow, op);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) bc_hubs_kernel(const int32_t* __restrict__ row, Op op,
                                                         const uint32_t* __restrict__ hubq,
                                                         const PlainCounters* c, unsigned cap,
                                                         const uint32_t* __restrict__ hub_off) {
  if (c->hub_count == 0) return;
  expand_hubs<kBlock>(row, hubq, c->hub_count, cap, hub_off, op);
}

__global__ void __launch_bounds__(kBlock) bc_fwd_expand_kernel(const int32_t* __restrict__ row, FwdOp op,
                                                               const uint32_t* __restrict__ q, unsigned qcount,
                                                               uint32_t* __restrict__ hubq, uint32_t hub) {
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  op.sq = &sq;
  expand_queue<kBlock>(row, q, qcount, hub, hubq, &op.c->hub_count, op.cap, &op.c->overflow, op);
  op.flush(true);
}

__global__ void __launch_bounds__(kBlock) bc_fwd_hubs_kernel(const int32_t* __restrict__ row, FwdOp op,
                                                             const uint32_t* __restrict__ hubq,
                                                             const uint32_t* __restrict__ hub_off) {
  if (op.c->hub_count == 0) return;
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  op.sq = &sq;
  expand_hubs<kBlock>(row, hubq, op.c->hub_count, op.cap, hub_off, op);
  op.flush(true);
}

__global__ void bc_plain_finish_level_kernel(const uint32_t* __restrict__ q, unsigned cnt,
                                             const float* __restrict__ sigma, double* __restrict__ delta,
                                             double* __restrict__ bc) {
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t v = q[i];
    const double d = delta[v] * static_cast<double>(sigma[v]);
    delta[v] = d;
    bc[v] += d;
  }
}

__global__ void bc_plain_reset_kernel(const uint32_t* __restrict__ q, unsigned cnt, uint32_t* __restrict__ level,
                                      float* __restrict__ sigma, double* __restrict__ delta) {
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t v = q[i];
    level[v] = kInf;
    sigma[v] = 0.f;
    delta[v] = 0.0;
  }
}

// ================================================================ batched
// Work item = (chunk, group); chunk = a vertex and at most `chunk` of its
// edges (in-edges for the forward pull, out-edges for the backward pull).
// The chunk length is the degree bin boundary of the batched sweeps: an
// adjacency up to one chunk is one wavefront's whole item, a longer one is
// split over several wavefronts that combine with atomics (knob bc.chunk,
// default kEdgeChunk).
// Forward levels per captured replay of the batched BC.
constexpr int kFwdLevels = 8;

struct ChunkTable {
  uint32_t* vtx;     // owning vertex
  int32_t* begin;    // first edge (end = min(begin + chunk, row[vtx + 1]))
  uint8_t* whole;    // 1 when the chunk is the vertex's whole adjacency
  int64_t count;
  uint32_t* multi;   // vertices with more than one chunk
  int64_t nmulti;
  int chunk;         // edges per chunk
};

__global__ void chunk_count_kernel(const int32_t* __restrict__ row, int64_t n, int chunk,
                                   uint32_t* __restrict__ cnt, uint32_t* __restrict__ is_multi) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t d = row[v + 1] - row[v];
    const uint32_t c = static_cast<uint32_t>((d + chunk - 1) / chunk);
    cnt[v] = c;
    is_multi[v] = c > 1 ? 1u : 0u;
  }
}

__global__ void chunk_fill_kernel(const int32_t* __restrict__ row, int64_t n, int chunk,
                                  const uint32_t* __restrict__ first,
                                  uint32_t* __restrict__ vtx, int32_t* __restrict__ begin,
                                  uint8_t* __restrict__ whole) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t b = row[v], e = row[v + 1];
    uint32_t k = first[v];
    const uint8_t w = (e - b) <= chunk ? 1 : 0;
    for (int32_t x = b; x < e; x += chunk, ++k) {
      vtx[k] = static_cast<uint32_t>(v);
      begin[k] = x;
      whole[k] = w;
    }
  }
}

__global__ void compact_flag_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int64_t n,
                                    uint32_t* __restrict__ out) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x)
    if (flag[v]) out[pos[v]] = static_cast<uint32_t>(v);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(x), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(x >> 32), lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

struct BatchDims {
  int64_t n;
  int groups;
  uint64_t last_mask;  // valid lanes of the last group
};

__device__ __forceinline__ uint64_t group_mask(const BatchDims& D, int g) {
  return g == D.groups - 1 ? D.last_mask : ~0ull;
}

// Vertices a 64-source group can reach: [vlo, vhi), the span of its
// sources' connected components (the whole graph when components are not
// known), and the chunks of those vertices [clo, clo + (next istart -
// istart)).  istart is the group's first work item in the group-major item
// order; entry `groups` holds the total.
struct GroupRange {
  int64_t vlo, vhi, clo, istart;
};

// Group ranges staged in LDS by every workgroup of a sweep.
struct GroupLds {
  int64_t istart[65], clo[65];
};

__device__ __forceinline__ void load_ranges(const GroupRange* __restrict__ R, int groups, GroupLds& L) {
  for (int i = threadIdx.x; i <= groups; i += blockDim.x) {
    L.istart[i] = R[i].istart;
    L.clo[i] = R[i].clo;
  }
  __syncthreads();
}

// Group of work item `it`: the last g with istart[g] <= it.
__device__ __forceinline__ int item_group(const GroupLds& L, int groups, int64_t it) {
  int lo = 0, hi = groups - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.istart[mid] <= it) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One thread: the ranges of every group from its sources.  comp_lo/comp_hi:
// per vertex the smallest id of its component and per component label one
// past its largest id (null: whole graph).  first: first chunk of every
// vertex (n + 1 entries).
__global__ void bcb_ranges_kernel(const uint32_t* __restrict__ src, int count, int groups, int64_t n,
                                  const uint32_t* __restrict__ comp_lo, const uint32_t* __restrict__ comp_hi,
                                  const uint32_t* __restrict__ first, GroupRange* __restrict__ R) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t acc = 0;
  for (int g = 0; g < groups; ++g) {
    int64_t lo = 0, hi = n;
    if (comp_lo) {
      lo = n;
      hi = 0;
      const int e = min(count, (g + 1) * 64);
      for (int i = g * 64; i < e; ++i) {
        const uint32_t l = comp_lo[src[i]];
        lo = min(lo, static_cast<int64_t>(l));
        hi = max(hi, static_cast<int64_t>(comp_hi[l]));
      }
    }
    R[g].vlo = lo;
    R[g].vhi = hi;
    R[g].clo = first[lo];
    R[g].istart = acc;
    acc += static_cast<int64_t>(first[hi]) - first[lo];
  }
  R[groups].vlo = R[groups].vhi = R[groups].clo = 0;
  R[groups].istart = acc;
}

// comp_hi[label] = 1 + largest vertex with that component label.  Vertices
// are visited from the highest id down and an atomic is issued only when it
// would raise the stored value, so a giant component settles its maximum in
// its first wavefronts and the rest only read it (one word would otherwise
// take an atomic per vertex).  A wavefront whose live lanes share one label
// reduces first.
__global__ void bcb_comp_hi_kernel(const uint32_t* __restrict__ label, int64_t n, uint32_t* __restrict__ comp_hi) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x) + (threadIdx.x & ~(kWave - 1)); b < n; b += stride) {
    const int64_t i = b + lane_id();
    const bool live = i < n;
    const int64_t v = n - 1 - i;  // highest ids first
    const uint32_t l = live ? label[v] : 0xffffffffu;
    const uint32_t l0 = __shfl(l, 0);  // lane 0 is live whenever any lane is
    const uint64_t same = __ballot(live && l == l0), alive = __ballot(live);
    if (same == alive) {
      const uint32_t hi = wave_max(live ? static_cast<uint32_t>(v + 1) : 0u);
      if (lane_id() == 0 && comp_hi[l0] < hi) atomicMax(&comp_hi[l0], hi);
    } else if (live && comp_hi[l] < static_cast<uint32_t>(v + 1)) {
      atomicMax(&comp_hi[l], static_cast<uint32_t>(v + 1));
    }
  }
}

// Level-0 masks, visited masks and path counts of the sources; the component
// label of every lane's source (lanes past `count`: no component).
__global__ void bcb_seed_kernel(const uint32_t* __restrict__ src, int count, BatchDims D, uint64_t* __restrict__ F0,
                                uint64_t* __restrict__ vis, float* __restrict__ sigma,
                                const uint32_t* __restrict__ comp, uint32_t* __restrict__ slab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D.groups * 64) slab[i] = (i < count && comp) ? comp[src[i]] : 0xffffffffu;
  if (i >= count) return;
  const int g = i / 64, j = i % 64;
  const uint32_t s = src[i];
  atomicOr(reinterpret_cast<unsigned long long*>(&F0[g * D.n + s]), 1ull << j);
  atomicOr(reinterpret_cast<unsigned long long*>(&vis[g * D.n + s]), 1ull << j);
  sigma[(static_cast<int64_t>(g) * D.n + s) * 64 + j] = 1.f;
}

// Device state of the forward level loop.
struct FwdState {
  unsigned long long found;  // new (vertex, source) pairs of the level (forward count)
  unsigned long long exact;  // the same, counted once per vertex by the visit kernel
  unsigned level;            // current level L (masks F[L] -> F[L + 1])
  unsigned done;             // set when a level finds nothing new; deepest = level
};

// Workgroup-uniform: adds the per-wavefront counts `v` (held by lane 0 of
// each wavefront) into *out with one atomic per workgroup; the level
// counters are single words that every wavefront of the grid would
// otherwise hit once per level.
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long* out) {
  __shared__ unsigned long long s_v[kBlock / kWave];
  if (lane_id() == 0) s_v[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / kWave; ++i) t += s_v[i];
    if (t) atomicAdd(out, t);
  }
}

// Start of a level: clears the counters and the next level's masks.
__global__ void bcb_level_begin_kernel(FwdState* S, uint64_t* const* __restrict__ F, int64_t total) {
  if (S->done) return;
  uint64_t* Fn = F[S->level + 1];
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    Fn[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->found = 0;
    S->exact = 0;
  }
}

// End of a level: stop when it found nothing, else move to the next level.
__global__ void bcb_level_end_kernel(FwdState* S) {
  if (threadIdx.x != 0 || S->done) return;
  if (S->exact == 0) S->done = 1;
  else S->level += 1;
}

// Forward pull for level L -> L+1 (L from the state).  slab: component label
// of each lane's source, comp: component label of each vertex (null when
// components are not known: every lane of the group).
__global__ void __launch_bounds__(kBlock) bcb_forward_kernel(
    const int32_t* __restrict__ trow, const int32_t* __restrict__ tcol, ChunkTable T, BatchDims D,
    uint64_t* const* __restrict__ F, FwdState* __restrict__ S, const uint64_t* __restrict__ vis,
    float* __restrict__ sigma, const GroupRange* __restrict__ R, const uint32_t* __restrict__ comp,
    const uint32_t* __restrict__ slab) {
  if (S->done) return;
  const uint64_t* __restrict__ FL = F[S->level];
  uint64_t* __restrict__ Fn = F[S->level + 1];
  unsigned long long* newcount = &S->found;
  const unsigned lane = lane_id();
  const int64_t wave = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
  __shared__ GroupLds GL;
  load_ranges(R, D.groups, GL);
  const int64_t items = GL.istart[D.groups];
  unsigned long long found = 0;
  for (int64_t it = wave; it < items; it += nwaves) {
    const int g = item_group(GL, D.groups, it);
    const int64_t c = GL.clo[g] + (it - GL.istart[g]);
    const uint32_t v = T.vtx[c];
    const int64_t goff = static_cast<int64_t>(g) * D.n;
    uint64_t lanes = group_mask(D, g);
    if (comp) lanes &= __ballot(slab[g * 64 + lane] == comp[v]);
    const uint64_t need = lanes & ~vis[goff + v];
    if (need == 0) continue;
    const int32_t eb = T.begin[c];
    const int32_t ee = min(eb + T.chunk, trow[v + 1]);
    uint64_t acc_mask = 0;
    float sig = 0.f;
    for (int32_t e0 = eb; e0 < ee; e0 += kWave) {
      const int32_t e = e0 + static_cast<int32_t>(lane);
      uint32_t u = 0;
      uint64_t f = 0;
      if (e < ee) {
        u = static_cast<uint32_t>(tcol[e]);
        f = FL[goff + u] & need;
      }
      uint64_t act = __ballot(f != 0);
      while (act) {
        const int i = __ffsll(static_cast<unsigned long long>(act)) - 1;
        act &= act - 1;
        const uint64_t fi = readlane64(f, i);
        const uint32_t ui = __builtin_amdgcn_readlane(u, i);
        acc_mask |= fi;
        if ((fi >> lane) & 1ull) sig += sigma[(goff + ui) * 64 + lane];
      }
    }
    if (acc_mask == 0) continue;
    float* sv = &sigma[(goff + v) * 64 + lane];
    if (T.whole[c]) {
      if ((acc_mask >> lane) & 1ull) *sv = sig;
      if (lane == 0) Fn[goff + v] = acc_mask;
    } else {
      if ((acc_mask >> lane) & 1ull) atomicAdd(sv, sig);
      if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(&Fn[goff + v]), acc_mask);
    }
    if (lane == 0) found += __popcll(acc_mask);
  }
  block_add(found, newcount);
}

// visited |= new frontier.  Counting of new (vertex, source) pairs is done
// by the forward kernel; multi-chunk vertices may be counted more than once
// there, so the exact count is recomputed here.
__global__ void __launch_bounds__(kBlock) bcb_visit_kernel(uint64_t* const* __restrict__ F, FwdState* S,
                                                           uint64_t* __restrict__ vis, int64_t total) {
  if (S->done) return;
  const uint64_t* __restrict__ Fn = F[S->level + 1];
  unsigned long long* exact = &S->exact;
  unsigned long long cnt = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint64_t f = Fn[i];
    if (f) {
      vis[i] |= f;
      cnt += __popcll(f);
    }
  }
  block_add(wave_sum(cnt), exact);
}

// Backward pull for level L (successors at L+1).  Whole-adjacency chunks
// finalise delta and the per-group bc partial directly; split adjacencies
// accumulate into delta and are finalised by bcb_multi_finish_kernel.
__global__ void __launch_bounds__(kBlock) bcb_backward_kernel(
    const int32_t* __restrict__ row, const int32_t* __restrict__ col, ChunkTable T, BatchDims D,
    const uint64_t* __restrict__ FL, const uint64_t* __restrict__ FL1, const float* __restrict__ sigma,
    double* __restrict__ delta, double* __restrict__ bcpart, const GroupRange* __restrict__ R) {
  const unsigned lane = lane_id();
  const int64_t wave = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
  __shared__ GroupLds GL;
  load_ranges(R, D.groups, GL);
  const int64_t items = GL.istart[D.groups];
  for (int64_t it = wave; it < items; it += nwaves) {
    const int g = item_group(GL, D.groups, it);
    const int64_t c = GL.clo[g] + (it - GL.istart[g]);
    const uint32_t v = T.vtx[c];
    const int64_t goff = static_cast<int64_t>(g) * D.n;
    const uint64_t mine = FL[goff + v];
    if (mine == 0) continue;
    const int32_t eb = T.begin[c];
    const int32_t ee = min(eb + T.chunk, row[v + 1]);
    double acc = 0.0;
    for (int32_t e0 = eb; e0 < ee; e0 += kWave) {
      const int32_t e = e0 + static_cast<int32_t>(lane);
      uint32_t w = 0;
      uint64_t f = 0;
      if (e < ee) {
        w = static_cast<uint32_t>(col[e]);
        f = FL1[goff + w] & mine;
      }
      uint64_t act = __ballot(f != 0);
      while (act) {
        const int i = __ffsll(static_cast<unsigned long long>(act)) - 1;
        act &= act - 1;
        const uint64_t fi = readlane64(f, i);
        const uint32_t wi = __builtin_amdgcn_readlane(w, i);
        if ((fi >> lane) & 1ull) {
          const int64_t idx = (goff + wi) * 64 + lane;
          acc += (1.0 + delta[idx]) / static_cast<double>(sigma[idx]);
        }
      }
    }
    const bool mylane = (mine >> lane) & 1ull;
    const int64_t vidx = (goff + v) * 64 + lane;
    if (T.whole[c]) {
      double d = 0.0;
      if (mylane) {
        d = acc * static_cast<double>(sigma[vidx]);
        delta[vidx] = d;
      }
      d = wave_sum(d);
      if (lane == 0) bcpart[goff + v] += d;
    } else if (mylane && acc != 0.0) {
      atomicAdd(&delta[vidx], acc);
    }
  }
}

// Split adjacencies at level L: delta = sigma * accumulated sum; bc partial.
__global__ void __launch_bounds__(kBlock) bcb_multi_finish_kernel(const uint32_t* __restrict__ multi,
                                                                  int64_t nmulti, BatchDims D,
                                                                  const uint64_t* __restrict__ FL,
                                                                  const float* __restrict__ sigma,
                                                                  double* __restrict__ delta,
                                                                  double* __restrict__ bcpart,
                                                                  const GroupRange* __restrict__ R) {
  const unsigned lane = lane_id();
  const int64_t wave = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
  for (int64_t it = wave; it < nmulti * D.groups; it += nwaves) {
    const int g = static_cast<int>(it / nmulti);
    const uint32_t v = multi[it - static_cast<int64_t>(g) * nmulti];
    if (v < R[g].vlo || v >= R[g].vhi) continue;
    const int64_t goff = static_cast<int64_t>(g) * D.n;
    const uint64_t mine = FL[goff + v];
    if (mine == 0) continue;
    double d = 0.0;
    if ((mine >> lane) & 1ull) {
      const int64_t idx = (goff + v) * 64 + lane;
      d = delta[idx] * static_cast<double>(sigma[idx]);
      delta[idx] = d;
    }
    d = wave_sum(d);
    if (lane == 0) bcpart[goff + v] += d;
  }
}

// bc[v] += sum over groups of the group partials, in group order.
__global__ void bcb_accumulate_kernel(const double* __restrict__ bcpart, BatchDims D, double* __restrict__ bc) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < D.n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < D.groups; ++g) s += bcpart[static_cast<int64_t>(g) * D.n + v];
    bc[v] += s;
  }
}

template <class T>
void fill_async(T* p, T v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  fill_kernel<T><<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(p, v, n);
  GX_LAUNCHED("fill_kernel");
}

// Chunk table of a CSR, cached in `buf`.
struct CachedChunks {
  DevBuf buf, first;  // first: first chunk of every vertex (n + 1)
  ChunkTable t{};
  bool ready = false;
};

void build_chunks(const int32_t* row, int64_t n, int64_t m, hipStream_t s, gx_graph* g, CachedChunks& cc,
                  int chunk) {
  if (cc.ready && cc.t.chunk == chunk) return;
  cc.ready = false;
  DevBuf cnt, flag, pos;
  DevBuf& first = cc.first;
  cnt.ensure((n + 1) * 4);
  first.ensure((n + 1) * 4);
  flag.ensure((n + 1) * 4);
  pos.ensure((n + 1) * 4);
  chunk_count_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(row, n, chunk, cnt.as<uint32_t>(),
                                                                    flag.as<uint32_t>());
  GX_LAUNCHED("chunk_count_kernel");
  fill_async<uint32_t>(cnt.as<uint32_t>() + n, 0u, 1, s);
  fill_async<uint32_t>(flag.as<uint32_t>() + n, 0u, 1, s);
  exclusive_scan_u32(cnt.as<uint32_t>(), first.as<uint32_t>(), n + 1, s);
  exclusive_scan_u32(flag.as<uint32_t>(), pos.as<uint32_t>(), n + 1, s);
  uint32_t nchunks = 0, nmulti = 0;
  read_bytes(g, &nchunks, first.as<uint32_t>() + n, 4);
  read_bytes(g, &nmulti, pos.as<uint32_t>() + n, 4);
  const int64_t C = nchunks, M = nmulti;
  // Layout: vtx[C] | begin[C] | multi[M] | whole[C] (bytes).
  cc.buf.ensure(C * 4 + C * 4 + M * 4 + C + 64);
  char* base = cc.buf.as<char>();
  cc.t.vtx = reinterpret_cast<uint32_t*>(base);
  cc.t.begin = reinterpret_cast<int32_t*>(base + C * 4);
  cc.t.multi = reinterpret_cast<uint32_t*>(base + C * 8);
  cc.t.whole = reinterpret_cast<uint8_t*>(base + C * 8 + M * 4);
  cc.t.count = C;
  cc.t.nmulti = M;
  cc.t.chunk = chunk;
  if (C > 0) {
    chunk_fill_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(row, n, chunk, first.as<uint32_t>(), cc.t.vtx,
                                                                    cc.t.begin, cc.t.whole);
    GX_LAUNCHED("chunk_fill_kernel");
  }
  if (M > 0) {
    compact_flag_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(flag.as<uint32_t>(), pos.as<uint32_t>(), n,
                                                                      cc.t.multi);
    GX_LAUNCHED("compact_flag_kernel");
  }
  GX_HIP(hipStreamSynchronize(s));
  (void)m;
  cc.ready = true;
}

// Per-graph caches of the batched BC (chunk tables of CSR and CSC).
struct BcCache {
  CachedChunks out, in;
  CachedChunks fixed;          // out-edge chunks of kEdgeChunk edges (out_edge_chunks)
  std::vector<DevBuf> levels;  // frontier masks per level
  DevBuf level_table;          // device copy of the level buffer pointers
  size_t table_cap = 0;
  CapturedBatch fwd;           // captured forward levels
  // Symmetric graphs: component label (smallest vertex id) per vertex and,
  // per label, one past the component's largest vertex id.
  DevBuf comp_lo, comp_hi;
  bool comps = false;
};

}  // namespace

// Graph-owned cache, created on first use, destroyed with the graph.
struct gx_bc_cache_holder {
  BcCache c;
};

static BcCache& bc_cache(gx_graph* g) {
  if (!g->bc_cache) g->bc_cache = std::make_shared<gx_bc_cache_holder>();
  return std::static_pointer_cast<gx_bc_cache_holder>(g->bc_cache)->c;
}

namespace gx {
EdgeChunks out_edge_chunks(gx_graph* g) {
  BcCache& cache = bc_cache(g);
  build_chunks(g->row, g->n, g->m, g->stream, g, cache.fixed, kEdgeChunk);
  return EdgeChunks{cache.fixed.t.vtx, cache.fixed.t.begin, cache.fixed.t.count};
}
}  // namespace gx

extern "C" gx_status gx_bc(gx_graph* g, const uint32_t* sources, uint32_t count, double* bc, gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(bc, "bc");
  GX_REQUIRE(count == 0 || sources != nullptr, "sources is NULL");
  for (uint32_t i = 0; i < count; ++i) check_vertex(g, sources[i], "source");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(std::max<int64_t>(n, 1));
  uint32_t* level = g->scratch<uint32_t>(WS_BC_LEVEL, cap);
  float* sigma = g->scratch<float>(WS_BC_SIGMA, cap);
  double* delta = g->scratch<double>(WS_BC_DELTA, cap);
  uint32_t* order = g->scratch<uint32_t>(WS_Q0, cap);
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, cap);
  uint32_t* hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(cap));
  PlainCounters* c = g->scratch<PlainCounters>(WS_COUNTERS, 1);
  PlainCounters* host = static_cast<PlainCounters*>(g->host_pinned(sizeof(PlainCounters)));
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0, levels_total = 0;
  uint64_t edges = 0;
  fill_async<double>(bc, 0.0, n, s);
  bc_plain_init_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(level, sigma, delta, n);
  GX_LAUNCHED("bc_plain_init_kernel");
  launches += 2;
  const unsigned grid = 256 * 8;
  const uint32_t hub = hub_degree();
  // In-edges for the path-count pull (the CSR itself for symmetric graphs).
  const int32_t* irow = g->row;
  const int32_t* icol = g->col;
  if (!(g->flags & GX_GRAPH_SYMMETRIC)) {
    build_transpose(g);
    irow = g->trow();
    icol = g->tcol();
  }
  std::vector<uint32_t> lstart;
  for (uint32_t si = 0; si < count; ++si) {
    const uint32_t src = sources[si];
    bc_plain_source_kernel<<<1, 64, 0, s>>>(level, sigma, order, c, src);
    lstart.assign({0u, 1u});
    uint32_t L = 0;
    // Forward: level L's vertices are order[lstart[L], lstart[L+1]).  Each
    // level is a discovery push followed by the path-count pull of the new
    // level.
    ktimer.begin();
    for (;;) {
      const uint32_t b = lstart[L], e = lstart[L + 1];
      const unsigned cnt = e - b;
      if (cnt == 0) break;
      bc_plain_clear_kernel<<<1, 64, 0, s>>>(c);
      FwdOp op{g->col, level, order + e, c, nullptr, L + 1, cap - e};
      bc_fwd_expand_kernel<<<std::min<unsigned>(grid, grid_for(cnt, kBlock)), kBlock, 0, s>>>(g->row, op,
                                                                                               order + b, cnt, hubq, hub);
      // The forward hub list is bounded by the level list's free room
      // (op.cap), as expand_hubs reads it.
      launch_hub_offsets<kBlock>(g->row, hubq, &c->hub_count, op.cap, hub_off, s);
      bc_fwd_hubs_kernel<<<grid, kBlock, 0, s>>>(g->row, op, hubq, hub_off);
      GX_LAUNCHED("bc forward");
      launches += 3;
      GX_HIP(hipMemcpyAsync(host, c, sizeof(PlainCounters), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
      if (host->overflow) fail(GX_ERR_CAPACITY, "BC level list overflow");
      const unsigned ncnt = host->count;
      lstart.push_back(e + ncnt);
      ++L;
      if (ncnt > 0) {
        bc_plain_clear_kernel<<<1, 64, 0, s>>>(c);
        SigmaOp sop{icol, level, sigma, L - 1};
        bc_expand_kernel<SigmaOp><<<std::min<unsigned>(grid, grid_for(ncnt, kBlock)), kBlock, 0, s>>>(
            irow, sop, order + e, ncnt, hubq, c, cap, hub);
        launch_hub_offsets<kBlock>(irow, hubq, &c->hub_count, cap, hub_off, s);
        bc_hubs_kernel<SigmaOp><<<grid, kBlock, 0, s>>>(irow, sop, hubq, c, cap, hub_off);
        GX_LAUNCHED("bc sigma");
        launches += 3;
      }
    }
    // Levels 0..L-1 are non-empty.  Backward from L-2 down to 1.
    const uint32_t deepest = L - 1;
    for (int64_t lv = static_cast<int64_t>(deepest) - 1; lv >= 1; --lv) {
      const uint32_t b = lstart[lv], e = lstart[lv + 1];
      const unsigned cnt = e - b;
      bc_plain_clear_kernel<<<1, 64, 0, s>>>(c);
      BwdOp op{g->col, level, sigma, delta, static_cast<uint32_t>(lv + 1)};
      bc_expand_kernel<BwdOp><<<std::min<unsigned>(grid, grid_for(cnt, kBlock)), kBlock, 0, s>>>(
          g->row, op, order + b, cnt, hubq, c, cap, hub);
      launch_hub_offsets<kBlock>(g->row, hubq, &c->hub_count, cap, hub_off, s);
      bc_hubs_kernel<BwdOp><<<grid, kBlock, 0, s>>>(g->row, op, hubq, c, cap, hub_off);
      bc_plain_finish_level_kernel<<<grid_for(cnt, kBlock, 4096), kBlock, 0, s>>>(order + b, cnt, sigma, delta, bc);
      GX_LAUNCHED("bc backward");
      launches += 4;
    }
    ktimer.end();
    const uint32_t total = lstart[L];
    bc_plain_reset_kernel<<<grid_for(total, kBlock, 4096), kBlock, 0, s>>>(order, total, level, sigma, delta);
    GX_LAUNCHED("bc_plain_reset_kernel");
    launches += 2;
    levels_total += L;
    edges += 0;
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = levels_total;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(g->m) * count;
    st->items = count;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}

extern "C" gx_status gx_bc_batched(gx_graph* g, const uint32_t* sources, uint32_t count, double* bc,
                                   gx_stats* st) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(bc, "bc");
  GX_REQUIRE(count == 0 || sources != nullptr, "sources is NULL");
  for (uint32_t i = 0; i < count; ++i) check_vertex(g, sources[i], "source");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  BcCache& cache = bc_cache(g);
  const int chunk = static_cast<int>(tune_effective("bc.chunk", n, g->m));
  const unsigned wgrid = static_cast<unsigned>(tune_effective("bc.grid", n, g->m));  // workgroups per sweep
  build_chunks(g->row, n, g->m, s, g, cache.out, chunk);
  const int32_t* trow = g->row;
  const int32_t* tcol = g->col;
  ChunkTable tin = cache.out.t;
  if (!(g->flags & GX_GRAPH_SYMMETRIC)) {
    build_transpose(g);
    trow = g->trow();
    tcol = g->tcol();
    build_chunks(trow, n, g->m, s, g, cache.in, chunk);
    tin = cache.in.t;
  }
  const ChunkTable tout = cache.out.t;
  const CachedChunks& cin = (g->flags & GX_GRAPH_SYMMETRIC) ? cache.out : cache.in;
  // Component bounds (symmetric graphs): a group only sweeps the span of its
  // sources' components, which matters when the graph is a union of many
  // (a batch of uneven graphs); one giant component leaves the whole graph.
  if ((g->flags & GX_GRAPH_SYMMETRIC) && !cache.comps && n > 0) {
    cache.comp_lo.ensure(n * 4);
    cache.comp_hi.ensure(n * 4);
    if (gx_cc(g, cache.comp_lo.as<uint32_t>(), nullptr) != GX_OK) fail(GX_ERR_HIP, "BC: component labels failed");
    fill_async<uint32_t>(cache.comp_hi.as<uint32_t>(), 0u, n, s);
    bcb_comp_hi_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(cache.comp_lo.as<uint32_t>(), n,
                                                                     cache.comp_hi.as<uint32_t>());
    GX_LAUNCHED("bcb_comp_hi_kernel");
    GX_HIP(hipStreamSynchronize(s));
    cache.comps = true;
  }
  const uint32_t* comp_lo = cache.comps ? cache.comp_lo.as<uint32_t>() : nullptr;
  const uint32_t* comp_hi = cache.comps ? cache.comp_hi.as<uint32_t>() : nullptr;
  // Groups per launch bounded by memory: sigma 256 B + delta 512 B + two
  // masks per vertex per group (plus the level masks).
  const int64_t per_group = n * (64 * 4 + 64 * 8 + 8 + 8);
  const int64_t budget = int64_t(48) << 30;
  int max_groups = static_cast<int>(std::max<int64_t>(1, budget / std::max<int64_t>(per_group, 1)));
  max_groups = std::min(max_groups, 64);
  RegionTimer region(s);
  KernelTimer ktimer(s);
  uint32_t launches = 0, levels_total = 0;
  fill_async<double>(bc, 0.0, n, s);
  ++launches;
  uint32_t* dsrc = g->scratch<uint32_t>(WS_Q3, std::max<uint32_t>(count, 1));
  if (count > 0) GX_HIP(hipMemcpyAsync(dsrc, sources, count * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  // Group ranges over the in-edge chunks (forward) and the out-edge chunks
  // (backward); 65 entries each.
  GroupRange* Rin = g->scratch<GroupRange>(WS_V5, 2 * 65);
  GroupRange* Rout = Rin + 65;
  std::vector<GroupRange> hr(65);
  for (uint32_t first = 0; first < count; first += 64u * max_groups) {
    const uint32_t cnt = std::min<uint32_t>(count - first, 64u * max_groups);
    BatchDims D;
    D.n = n;
    D.groups = static_cast<int>((cnt + 63) / 64);
    const uint32_t tail = cnt - 64u * (D.groups - 1);
    D.last_mask = tail == 64 ? ~0ull : ((1ull << tail) - 1);
    const int64_t gn = static_cast<int64_t>(D.groups) * n;
    float* sigma = g->scratch<float>(WS_BC_SIGMA, gn * 64);
    double* delta = g->scratch<double>(WS_BC_DELTA, gn * 64);
    uint64_t* vis = g->scratch<uint64_t>(WS_BC_MASK0, gn);
    double* bcpart = g->scratch<double>(WS_BC_MASK1, gn);
    bcb_ranges_kernel<<<1, 64, 0, s>>>(dsrc + first, static_cast<int>(cnt), D.groups, n, comp_lo, comp_hi,
                                         cin.first.as<uint32_t>(), Rin);
    bcb_ranges_kernel<<<1, 64, 0, s>>>(dsrc + first, static_cast<int>(cnt), D.groups, n, comp_lo, comp_hi,
                                         cache.out.first.as<uint32_t>(), Rout);
    GX_LAUNCHED("bcb_ranges_kernel");
    GX_HIP(hipMemcpyAsync(hr.data(), Rout, (D.groups + 1) * sizeof(GroupRange), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    launches += 2;
    // Path counts and dependencies are read only at reached vertices, so only
    // each group's span is cleared.
    for (int gi = 0; gi < D.groups; ++gi) {
      GX_REQUIRE(hr[gi].vlo >= 0 && hr[gi].vlo <= hr[gi].vhi && hr[gi].vhi <= n, "BC: bad group range");
      const int64_t off = (static_cast<int64_t>(gi) * n + hr[gi].vlo) * 64;
      fill_async<float>(sigma + off, 0.f, (hr[gi].vhi - hr[gi].vlo) * 64, s);
      fill_async<double>(delta + off, 0.0, (hr[gi].vhi - hr[gi].vlo) * 64, s);
      launches += 2;
    }
    fill_async<uint64_t>(vis, 0ull, gn, s);
    fill_async<double>(bcpart, 0.0, gn, s);
    launches += 2;
    auto level_buf = [&](size_t L) -> uint64_t* {
      while (cache.levels.size() <= L) cache.levels.emplace_back();
      cache.levels[L].ensure(gn * sizeof(uint64_t));
      return cache.levels[L].as<uint64_t>();
    };
    // Level buffers are allocated ahead of each replay of the captured
    // forward levels; the device table holds their addresses.
    auto ensure_levels = [&](size_t upto) {
      for (size_t L = 0; L <= upto; ++L) level_buf(L);
      if (cache.table_cap < cache.levels.size()) {
        cache.table_cap = std::max<size_t>(cache.levels.size(), 2 * cache.table_cap);
        cache.level_table.ensure(cache.table_cap * sizeof(uint64_t*));
      }
      std::vector<uint64_t*> ptrs(cache.levels.size());
      for (size_t L = 0; L < ptrs.size(); ++L) ptrs[L] = cache.levels[L].as<uint64_t>();
      GX_HIP(hipMemcpyAsync(cache.level_table.p, ptrs.data(), ptrs.size() * sizeof(uint64_t*),
                            hipMemcpyHostToDevice, s));
      GX_HIP(hipStreamSynchronize(s));
    };
    ensure_levels(kFwdLevels + 1);
    uint64_t* F0 = cache.levels[0].as<uint64_t>();
    fill_async<uint64_t>(F0, 0ull, gn, s);
    uint32_t* slab = g->scratch<uint32_t>(WS_V4, 64 * 64);
    bcb_seed_kernel<<<grid_for(D.groups * 64, kBlock), kBlock, 0, s>>>(dsrc + first, static_cast<int>(cnt), D, F0,
                                                                        vis, sigma, comp_lo, slab);
    GX_LAUNCHED("bcb_seed_kernel");
    FwdState* fst = g->scratch<FwdState>(WS_BC_LEVEL, 1);
    fill_async<unsigned long long>(reinterpret_cast<unsigned long long*>(fst), 0ull,
                                   sizeof(FwdState) / sizeof(unsigned long long), s);
    launches += 3;
    // Forward sweep: level masks F[0..deepest], kFwdLevels levels per replay.
    ktimer.begin();
    FwdState fh{};
    {
      uint64_t key = 0;
      for (const void* p : {(const void*)trow, (const void*)tcol, (const void*)tin.vtx, (const void*)vis,
                            (const void*)sigma, (const void*)Rin, (const void*)comp_lo, (const void*)slab,
                            (const void*)fst})
        key = mix_key(key, reinterpret_cast<uintptr_t>(p));
      key = mix_key(mix_key(mix_key(key, static_cast<uint64_t>(gn)), D.last_mask), cache.table_cap);
      key = mix_key(mix_key(mix_key(key, static_cast<uint64_t>(tin.count)), static_cast<uint64_t>(tin.chunk)), wgrid);
      uint64_t* const* table = cache.level_table.as<uint64_t*>();
      for (;;) {
        const uint64_t k = mix_key(key, reinterpret_cast<uintptr_t>(cache.level_table.p));
        g->replay(cache.fwd, k, [&](hipStream_t cs) {
          for (int i = 0; i < kFwdLevels; ++i) {
            bcb_level_begin_kernel<<<grid_for(gn, kBlock, 8192), kBlock, 0, cs>>>(fst, table, gn);
            bcb_forward_kernel<<<wgrid, kBlock, 0, cs>>>(trow, tcol, tin, D, table, fst, vis, sigma, Rin, comp_lo,
                                                         slab);
            bcb_visit_kernel<<<grid_for(gn, kBlock, 8192), kBlock, 0, cs>>>(table, fst, vis, gn);
            bcb_level_end_kernel<<<1, 64, 0, cs>>>(fst);
          }
          GX_LAUNCHED("bcb forward levels");
        });
        launches += 4 * kFwdLevels;
        GX_HIP(hipMemcpyAsync(&fh, fst, sizeof(FwdState), hipMemcpyDeviceToHost, s));
        GX_HIP(hipStreamSynchronize(s));
        if (fh.done) break;
        GX_REQUIRE(fh.level < 65535 - kFwdLevels, "BC level count exceeds the level table");
        const size_t before = cache.table_cap;
        ensure_levels(fh.level + kFwdLevels + 1);
        table = cache.level_table.as<uint64_t*>();
        if (cache.table_cap != before) key = mix_key(key, cache.table_cap);
      }
    }
    size_t L = fh.level;
    // Levels 0..L are non-empty (L = deepest).  Backward sweep L-1 .. 1,
    // captured once into a HIP graph and replayed.
    const size_t deepest = L;
    levels_total += static_cast<uint32_t>(deepest + 1);
    if (deepest >= 2) {
      hipGraph_t graph;
      hipGraphExec_t exec;
      // Capture on the graph's own non-null stream (the null stream cannot be
      // captured); the forward work on `s` has completed (synchronised above).
      hipStream_t cs = g->capture();
      GX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      for (int64_t lv = static_cast<int64_t>(deepest) - 1; lv >= 1; --lv) {
        const uint64_t* FL = cache.levels[lv].as<uint64_t>();
        const uint64_t* FL1 = cache.levels[lv + 1].as<uint64_t>();
        bcb_backward_kernel<<<wgrid, kBlock, 0, cs>>>(g->row, g->col, tout, D, FL, FL1, sigma, delta, bcpart,
                                                       Rout);
        if (tout.nmulti > 0)
          bcb_multi_finish_kernel<<<grid_for(tout.nmulti * D.groups * 64, kBlock, 8192), kBlock, 0, cs>>>(
              tout.multi, tout.nmulti, D, FL, sigma, delta, bcpart, Rout);
      }
      hipError_t cap_err = hipStreamEndCapture(cs, &graph);
      GX_HIP(cap_err);
      GX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      GX_HIP(hipGraphLaunch(exec, s));
      GX_HIP(hipStreamSynchronize(s));
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      launches += static_cast<uint32_t>((deepest - 1) * (tout.nmulti > 0 ? 2 : 1));
    }
    ktimer.end();
    bcb_accumulate_kernel<<<grid_for(n, kBlock, 8192), kBlock, 0, s>>>(bcpart, D, bc);
    GX_LAUNCHED("bcb_accumulate_kernel");
    ++launches;
  }
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = levels_total;
    st->launches = launches;
    st->edges_visited = static_cast<uint64_t>(g->m) * count;
    st->items = count;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
