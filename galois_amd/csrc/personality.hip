// D-IrGL personality boundary: the per-operator GPU kernels that D-IrGL host
// code launches on one host's partition.
//
// Reference
//  - context and graph: get_CUDA_context / init_CUDA_context /
//    load_graph_CUDA (libgluon/include/galois/cuda/HostDecls.h:38-54,98-102);
//  - BFS push operators InitializeGraph, FirstItr_BFS, BFS, BFSSanityCheck
//    (lonestar/analytics/distributed/bfs/bfs_push.cpp:95-340, declared in
//    bfs_push_cuda.h:103-140) over the fields dist_current / dist_old;
//  - SSSP push (sssp/sssp_push.cpp:95-360, sssp_push_cuda.h:103-128): BFS
//    push with dist_current + edge weight, and a distance sum in the check;
//  - CC push (connected-components/cc_push.cpp:85-300, cc_push_cuda.h:
//    103-130) over comp_current / comp_old: labels start at the global id,
//    every node with a lowered label pushes it, active counted per edge;
//  - PageRank push (pagerank/pagerank_push.cpp:110-430, pagerank_push_cuda.h)
//    over value / nout / residual / delta (FP32, stored as bit patterns):
//    residual flows to value, deltas above the tolerance are pushed with
//    float atomics (order dependent, as the reference's);
//  - BC level (betweennesscentrality/bc_level.cpp:150-560, bc_level_cuda.h:
//    223-275) over current_length, num_shortest_paths (FP64),
//    dependency and betweeness_centrality (FP32): level-synchronous forward
//    passes counting shortest paths, backward passes accumulating
//    dependencies one node per wavefront (fixed summation order);
//  - k-core push operators InitializeGraph1/2, KCoreStep1/2, KCoreSanityCheck
//    (lonestar/analytics/distributed/k-core/kcore_push.cpp:93-360, declared
//    in kcore_push_cuda.h:133-167) over current_degree / trim / flag;
//  - field accessors get/set/add/min_node_<field>_cuda, batch_reset, and the
//    change bitsets get_bitset_<field>_cuda / bitset_<field>_reset_cuda.
//
// Every operator runs over local ids [begin, end) (the _allNodes /
// _masterNodes / _nodesWithEdges variants are three such ranges, see
// gx_part_range) and returns its reductions through out-parameters, as the
// reference's DGAccumulator arguments.  Operators that walk edges use the
// load-balanced expansion of frontier.hpp with the range as an identity
// frontier and the operator's vertex test as the admission filter, so hub
// vertices are spread over the grid instead of serialising one thread.
// The multi-host mirror exchange (batch_get/set/min over master/mirror
// lists) belongs to gluon and is not part of this one-GPU build.
#include <algorithm>
#include <cstring>
#include <vector>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxGrid = 2048;
constexpr int kFields = 15;  // field GX_FIELD_BC_PATHS holds doubles

// Device counters of one operator launch.
struct PartCounters {
  unsigned hub_count;
  unsigned overflow;
  unsigned active;
  unsigned changed;
  unsigned long long work;
  unsigned long long sum;
  unsigned long long dsum;
  unsigned max;
  unsigned pad;
  // PageRank sanity (non-negative floats compared as bit patterns)
  unsigned max_value, min_value, max_res, min_res;
  double sum_value, sum_res;
  unsigned long long over_tol;
};

}  // namespace

struct gx_part {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t n = 0, m = 0;
  uint32_t num_owned = 0, begin_master = 0, with_edges = 0;
  bool loaded = false;
  gx::DevBuf row, col, w, gid, field[kFields], bits[kFields], hubq, hub_off, counters, push, bq;
  PartCounters* host = nullptr;  // pinned
  ~gx_part() {
    if (host) (void)hipHostFree(host);
    if (stream) (void)hipStreamDestroy(stream);
  }
  uint32_t* f(int k) const { return field[k].as<uint32_t>(); }
  unsigned long long* b(int k) const { return bits[k].as<unsigned long long>(); }
  int64_t words() const { return (n + 63) / 64; }
};

namespace {

void check_part(const gx_part* p) {
  GX_REQUIRE(p != nullptr, "partition handle is NULL");
  GX_REQUIRE(p->loaded, "partition has no graph (gx_part_load)");
}
// Fields read and written as uint32 words (all but the FP64 path counts).
void check_field(int k) {
  GX_REQUIRE(k >= 0 && k < kFields, "unknown node field");
  GX_REQUIRE(k != GX_FIELD_BC_PATHS, "num_shortest_paths is FP64: use gx_part_read_paths");
}
void check_bitset_field(int k) { GX_REQUIRE(k >= 0 && k < kFields, "unknown node field"); }
void check_range(const gx_part* p, uint32_t begin, uint32_t end) {
  GX_REQUIRE(begin <= end && static_cast<int64_t>(end) <= p->n, "range [begin, end) outside the partition");
}

// Marks v in a change bitset.  A plain load first: once a word's bit is set
// the atomic is skipped, so hot words (64 neighbouring nodes, pushed to by
// every in-edge) do not serialise on atomics.
__device__ __forceinline__ void set_bit(unsigned long long* bits, uint32_t v) {
  const unsigned long long m = 1ull << (v & 63);
  if (!(__atomic_load_n(&bits[v >> 6], __ATOMIC_RELAXED) & m)) atomicOr(&bits[v >> 6], m);
}

// Per-lane counters reduced once per wavefront at the end of a launch.
__device__ __forceinline__ void flush_counts(PartCounters* c, unsigned active, unsigned long long work) {
  active = wave_sum(active);
  work = wave_sum(work);
  if (lane_id() == 0) {
    if (active) atomicAdd(&c->active, active);
    if (work) atomicAdd(&c->work, work);
  }
}

__global__ void fill_kernel_u32(uint32_t* __restrict__ f, uint32_t begin, uint32_t end, uint32_t v) {
  for (uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x) f[i] = v;
}

// Clears the bits of [begin, end): whole words stored, edge words masked.
__global__ void bitset_clear_kernel(unsigned long long* __restrict__ bits, uint32_t begin, uint32_t end) {
  const uint32_t w0 = begin >> 6, w1 = (end - 1) >> 6;
  for (uint32_t w = w0 + blockIdx.x * blockDim.x + threadIdx.x; w <= w1; w += gridDim.x * blockDim.x) {
    const uint32_t lo = max(begin, w << 6), hi = min(end, (w << 6) + 64);  // bits [lo, hi) of word w
    const unsigned long long mask =
        (hi - lo == 64) ? ~0ull : (((1ull << (hi - lo)) - 1) << (lo & 63));
    if (mask == ~0ull)
      bits[w] = 0;
    else
      atomicAnd(&bits[w], ~mask);
  }
}

__global__ void node_update_kernel(uint32_t* __restrict__ f, uint32_t lid, uint32_t v, int op,
                                   PartCounters* __restrict__ c) {
  if (threadIdx.x != 0) return;
  if (op == 0) {
    atomicAdd(&f[lid], v);
  } else {
    const uint32_t old = atomicMin(&f[lid], v);
    c->changed = old > v ? 1u : 0u;
  }
}

// ---------------------------------------------------------------- BFS push
// InitializeGraph (bfs_push.cpp:118-125): dist = 0 at the source's global
// id, infinity elsewhere.
__global__ void bfs_init_kernel(uint32_t* __restrict__ cur, uint32_t* __restrict__ old,
                                const uint32_t* __restrict__ gid, uint32_t begin, uint32_t end, uint32_t inf,
                                uint64_t src) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint64_t g = gid ? gid[v] : v;
    const uint32_t d = g == src ? 0u : inf;
    cur[v] = d;
    old[v] = d;
  }
}

// The min-push operators of BFS, SSSP and CC push.  FirstItr_* (bfs_push.cpp:
// 170-181, sssp_push.cpp:173-185, cc_push.cpp:152-164) admits its whole
// range; BFS / SSSP (bfs_push.cpp:254-273, sssp_push.cpp:265-286) admit
// nodes with dist_old > dist_current (counted active) whose dist_current is
// below the priority; ConnectedComp (cc_push.cpp:229-246) admits nodes with
// comp_old > comp_current and counts one active per pushed edge.  An admitted
// node copies its current value into old and pushes current + 1 (BFS),
// current + weight (SSSP) or current (CC) to every out-neighbour with
// atomicMin, marking lowered neighbours in the current field's bitset.
enum PushKind : int { kPushBfs = 0, kPushSssp = 1, kPushCc = 2 };

struct MinPushOp {
  const int32_t* __restrict__ col;
  const uint32_t* __restrict__ w;
  uint32_t* cur;
  uint32_t* old;
  unsigned long long* bits;
  PartCounters* c;
  uint32_t priority;
  int kind;
  bool first;
  unsigned active = 0;
  unsigned long long work = 0;
  __device__ __forceinline__ bool admit(uint32_t v) {
    const uint32_t dc = cur[v];
    if (first) {
      old[v] = dc;
      return true;
    }
    if (old[v] > dc) {
      if (kind == kPushCc) {
        old[v] = dc;
        return true;
      }
      ++active;
      if (priority > dc) {
        old[v] = dc;
        return true;
      }
    }
    return false;
  }
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid) return;
    const uint32_t v = static_cast<uint32_t>(col[e]);
    const uint32_t base = cur[u];
    const uint32_t nd = kind == kPushBfs ? base + 1u : (kind == kPushSssp ? base + w[e] : base);
    // Labels only decrease, so a plain read at or below nd (a stale one is
    // larger, never smaller) means the atomic could not lower the label:
    // hubs' labels are read by many pushers and settle early.
    if (nd < cur[v]) {
      const uint32_t prev = atomicMin(&cur[v], nd);
      if (prev > nd) set_bit(bits, v);
    }
    if (!first) {
      if (kind == kPushCc)
        ++active;
      else
        ++work;
    }
  }
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void finish() { flush_counts(c, active, work); }
};

// BFSSanityCheck / SSSPSanityCheck (bfs_push.cpp:323-330, sssp_push.cpp:
// 351-358): reached count, distance sum and maximum distance.
__global__ void __launch_bounds__(kBlock) dist_check_kernel(const uint32_t* __restrict__ cur, uint32_t begin,
                                                            uint32_t end, uint32_t inf, PartCounters* c) {
  unsigned long long cnt = 0, dsum = 0;
  uint32_t mx = 0;
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint32_t d = cur[v];
    if (d < inf) {
      ++cnt;
      dsum += d;
      mx = max(mx, d);
    }
  }
  cnt = wave_sum(cnt);
  dsum = wave_sum(dsum);
  mx = wave_max(mx);
  if (lane_id() == 0) {
    if (cnt) atomicAdd(&c->sum, cnt);
    if (dsum) atomicAdd(&c->dsum, dsum);
    atomicMax(&c->max, mx);
  }
}

// CC InitializeGraph (cc_push.cpp:109-113): both labels = the global id.
__global__ void cc_init_kernel(uint32_t* __restrict__ cur, uint32_t* __restrict__ old,
                               const uint32_t* __restrict__ gid, uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint32_t g = gid ? gid[v] : v;
    cur[v] = g;
    old[v] = g;
  }
}

// ConnectedCompSanityCheck (cc_push.cpp:290-296): nodes whose label is their
// own global id (one per component).
__global__ void __launch_bounds__(kBlock) cc_check_kernel(const uint32_t* __restrict__ cur,
                                                          const uint32_t* __restrict__ gid, uint32_t begin,
                                                          uint32_t end, PartCounters* c) {
  unsigned long long cnt = 0;
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x)
    cnt += cur[v] == (gid ? gid[v] : v) ? 1 : 0;
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(&c->sum, cnt);
}

// --------------------------------------------------------- PageRank push
__device__ __forceinline__ float as_f(uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ uint32_t as_u(float f) { return __float_as_uint(f); }

// ResetGraph (pagerank_push.cpp:118-124): all four fields zero.
__global__ void pr_reset_kernel(uint32_t* __restrict__ value, uint32_t* __restrict__ nout,
                                uint32_t* __restrict__ res, uint32_t* __restrict__ delta, uint32_t begin,
                                uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    value[v] = 0;
    nout[v] = 0;
    res[v] = 0;
    delta[v] = 0;
  }
}

// InitializeGraph (pagerank_push.cpp:168-175): residual = alpha, nout += the
// node's out-degree (bitset nout).
__global__ void pr_init_kernel(const int32_t* __restrict__ row, uint32_t* __restrict__ nout,
                               uint32_t* __restrict__ res, unsigned long long* __restrict__ nout_bits, float alpha,
                               uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    res[v] = as_u(alpha);
    atomicAdd(&nout[v], static_cast<uint32_t>(row[v + 1] - row[v]));
    set_bit(nout_bits, v);
  }
}

// PageRank_delta (pagerank_push.cpp:210-223): residual moves into value; a
// residual above the tolerance becomes the delta residual*(1-alpha)/nout.
__global__ void pr_delta_kernel(uint32_t* __restrict__ value, const uint32_t* __restrict__ nout,
                                uint32_t* __restrict__ res, uint32_t* __restrict__ delta, float alpha, float tol,
                                uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const float r = as_f(res[v]);
    if (r > 0.0f) {
      res[v] = 0;
      value[v] = as_u(as_f(value[v]) + r);
      const uint32_t k = nout[v];
      if (r > tol && k > 0) delta[v] = as_u(r * (1.0f - alpha) / static_cast<float>(k));
    }
  }
}

// PageRank (pagerank_push.cpp:290-307): a node with a positive delta (counted
// active) clears it and adds it to every out-neighbour's residual (float
// atomics, bitset residual).  The delta is parked in `push` for the edges.
// Pushes to the local ids below kHotIds (on generator graphs and degree-ordered
// ones, the hubs, whose residual words otherwise take one atomic per
// in-neighbour one after another) are summed per workgroup in LDS first and
// added with one atomic per id at the end of the launch; every push is
// positive, so an id was pushed to exactly when its LDS sum is nonzero.
constexpr uint32_t kHotIds = 4096;

struct PrPushOp {
  const int32_t* __restrict__ col;
  uint32_t* delta;
  float* push;
  float* res;
  unsigned long long* bits;
  PartCounters* c;
  unsigned active = 0;
  float* hot = nullptr;  // LDS sums of the pushes to ids below kHotIds
  __device__ __forceinline__ bool admit(uint32_t v) {
    const float d = as_f(delta[v]);
    if (!(d > 0.0f)) return false;
    delta[v] = 0;
    push[v] = d;
    ++active;
    return true;
  }
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid) return;
    const uint32_t v = static_cast<uint32_t>(col[e]);
    if (v < kHotIds) {
      atomicAdd(&hot[v], push[u]);
      return;
    }
    atomicAdd(&res[v], push[u]);
    set_bit(bits, v);
  }
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void finish() { flush_counts(c, active, 0ull); }
  // Workgroup-uniform.  Ids at or above n are never pushed to.
  __device__ __forceinline__ void flush_hot() {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHotIds; i += blockDim.x) {
      const float sum = hot[i];
      if (sum != 0.0f) {
        atomicAdd(&res[i], sum);
        set_bit(bits, i);
      }
    }
  }
};



// PageRankSanity (pagerank_push.cpp:415-429): value and residual extremes and
// sums, residuals above the tolerance.
__global__ void __launch_bounds__(kBlock) pr_check_kernel(const uint32_t* __restrict__ value,
                                                          const uint32_t* __restrict__ res, float tol,
                                                          uint32_t begin, uint32_t end, PartCounters* c) {
  uint32_t mxv = 0, mnv = 0x7f800000u, mxr = 0, mnr = 0x7f800000u;
  double sv = 0, sr = 0;
  unsigned long long over = 0;
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint32_t bv = value[v], br = res[v];
    mxv = max(mxv, bv);
    mnv = min(mnv, bv);
    mxr = max(mxr, br);
    mnr = min(mnr, br);
    sv += as_f(bv);
    sr += as_f(br);
    over += as_f(br) > tol ? 1 : 0;
  }
  mxv = wave_max(mxv);
  mnv = wave_min(mnv);
  mxr = wave_max(mxr);
  mnr = wave_min(mnr);
  sv = wave_sum(sv);
  sr = wave_sum(sr);
  over = wave_sum(over);
  if (lane_id() == 0) {
    atomicMax(&c->max_value, mxv);
    atomicMin(&c->min_value, mnv);
    atomicMax(&c->max_res, mxr);
    atomicMin(&c->min_res, mnr);
    atomicAdd(&c->sum_value, sv);
    atomicAdd(&c->sum_res, sr);
    if (over) atomicAdd(&c->over_tol, over);
  }
}

// ------------------------------------------------------------------ BC level
// InitializeGraph (bc_level.cpp:154-160): centrality, path counts and
// dependencies zero.
__global__ void bc_init_kernel(uint32_t* __restrict__ bc, uint32_t* __restrict__ dep, double* __restrict__ nsp,
                               uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    bc[v] = 0;
    dep[v] = 0;
    nsp[v] = 0.0;
  }
}

// InitializeIteration (bc_level.cpp:203-216): the source at length 0 with
// one path, every other node at infinity with none; dependencies zero.
__global__ void bc_iter_kernel(uint32_t* __restrict__ len, double* __restrict__ nsp, uint32_t* __restrict__ dep,
                               const uint32_t* __restrict__ gid, uint32_t begin, uint32_t end, uint32_t inf,
                               uint64_t src) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const bool is_src = (gid ? static_cast<uint64_t>(gid[v]) : static_cast<uint64_t>(v)) == src;
    len[v] = is_src ? 0u : inf;
    nsp[v] = is_src ? 1.0 : 0.0;
    dep[v] = 0;
  }
}

// ForwardPass (bc_level.cpp:290-322): nodes at length r push r + 1; a
// neighbour reached at r + 1 (first or again) adds the node's path count
// (bitsets current_length / num_shortest_paths), each counted in dga.
struct BcForwardOp {
  const int32_t* __restrict__ col;
  uint32_t* len;
  double* nsp;
  unsigned long long* len_bits;
  unsigned long long* nsp_bits;
  PartCounters* c;
  uint32_t r;
  unsigned dga = 0;
  __device__ __forceinline__ bool admit(uint32_t v) { return len[v] == r; }
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid) return;
    const uint32_t v = static_cast<uint32_t>(col[e]);
    const uint32_t nd = r + 1;
    if (len[v] < nd) return;  // reached at an earlier level (a stale read is larger, never smaller)
    const uint32_t old = atomicMin(&len[v], nd);
    if (old >= nd) {
      if (old > nd) set_bit(len_bits, v);
      atomicAdd(&nsp[v], nsp[u]);
      set_bit(nsp_bits, v);
      ++dga;
    }
  }
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void finish() { flush_counts(c, dga, 0ull); }
};

// MiddleSync (bc_level.cpp:368-374): reached nodes marked in the path-count
// bitset.
__global__ void bc_middle_kernel(const uint32_t* __restrict__ len, unsigned long long* __restrict__ nsp_bits,
                                 uint32_t begin, uint32_t end, uint32_t inf) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x)
    if (len[v] != inf) set_bit(nsp_bits, v);
}

// BackwardPass, part 1: the nodes of [begin, end) at length r, in a queue.
__global__ void __launch_bounds__(kBlock) bc_level_queue_kernel(const uint32_t* __restrict__ len, uint32_t begin,
                                                                uint32_t end, uint32_t r, uint32_t* __restrict__ q,
                                                                PartCounters* c) {
  __shared__ LdsQueue<1024> lq;
  lq_init(lq);
  const uint32_t n = end - begin;
  for (uint32_t b0 = blockIdx.x * kBlock; b0 < n; b0 += gridDim.x * kBlock) {
    const uint32_t i = b0 + threadIdx.x;
    const bool hit = i < n && len[begin + i] == r;
    lq_push(lq, hit, begin + i);
    lq_flush(lq, q, &c->active, n, &c->overflow, false);
  }
  lq_flush(lq, q, &c->active, n, &c->overflow, true);
}

// BackwardPass, part 2 (bc_level.cpp:415-433): one wavefront per queued node
// sums (1 + dependency) / paths over its successors at length r + 1 in a
// fixed tree order, then scales by the node's own path count; FP32 result
// as the reference's field (bitset dependency when a successor exists).
__global__ void __launch_bounds__(kBlock) bc_backward_kernel(const int32_t* __restrict__ row,
                                                             const int32_t* __restrict__ col,
                                                             const uint32_t* __restrict__ len,
                                                             const double* __restrict__ nsp,
                                                             uint32_t* __restrict__ dep,
                                                             unsigned long long* __restrict__ dep_bits,
                                                             const uint32_t* __restrict__ q, const PartCounters* c,
                                                             uint32_t r) {
  const unsigned count = c->active;
  const unsigned lane = lane_id();
  const unsigned waves = gridDim.x * (kBlock / kWave);
  for (unsigned i = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; i < count; i += waves) {
    const uint32_t u = q[i];
    double sum = 0.0;
    unsigned found = 0;
    for (int32_t e = row[u] + static_cast<int32_t>(lane); e < row[u + 1]; e += kWave) {
      const uint32_t v = static_cast<uint32_t>(col[e]);
      if (len[v] == r + 1) {
        sum += (1.0 + static_cast<double>(__uint_as_float(dep[v]))) / nsp[v];
        found = 1;
      }
    }
    sum = wave_sum(sum);
    found = wave_sum(found);
    if (lane == 0 && found) {
      const float d = static_cast<float>(static_cast<double>(__uint_as_float(dep[u])) + sum);
      dep[u] = __float_as_uint(static_cast<float>(static_cast<double>(d) * nsp[u]));
      set_bit(dep_bits, u);
    }
  }
}

// BC (bc_level.cpp:480-485): a positive dependency adds to the centrality.
__global__ void bc_accumulate_kernel(uint32_t* __restrict__ bc, const uint32_t* __restrict__ dep, uint32_t begin,
                                     uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const float d = __uint_as_float(dep[v]);
    if (d > 0.0f) bc[v] = __float_as_uint(__uint_as_float(bc[v]) + d);
  }
}

// Sanity (bc_level.cpp:525-545): sum, maximum and minimum centrality.
__global__ void __launch_bounds__(kBlock) bc_check_kernel(const uint32_t* __restrict__ bc, uint32_t begin,
                                                          uint32_t end, PartCounters* c) {
  uint32_t mx = 0, mn = 0x7f800000u;
  double sum = 0;
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint32_t b = bc[v];
    mx = max(mx, b);
    mn = min(mn, b);
    sum += __uint_as_float(b);
  }
  mx = wave_max(mx);
  mn = wave_min(mn);
  sum = wave_sum(sum);
  if (lane_id() == 0) {
    atomicMax(&c->max_value, mx);
    atomicMin(&c->min_value, mn);
    atomicAdd(&c->sum_value, sum);
  }
}

// ------------------------------------------------------------- k-core push
// InitializeGraph1 (kcore_push.cpp:140-181): flag set,
// trim and current_degree zero.
__global__ void kcore_init1_kernel(uint32_t* __restrict__ deg, uint32_t* __restrict__ trim,
                                   uint32_t* __restrict__ flag, uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    flag[v] = 1;
    trim[v] = 0;
    deg[v] = 0;
  }
}

// InitializeGraph2 (kcore_push.cpp:93-137: every out-edge adds one to the
// destination's degree) and KCoreStep1 (kcore_push.cpp:227-318: a live
// vertex below k dies, counted active, and adds one
// to the trim of every out-neighbour); both mark the bitset of the field
// they add to.
struct KcoreOp {
  const int32_t* __restrict__ col;
  uint32_t* deg;
  uint32_t* trim;
  uint32_t* flag;
  uint32_t* target;  // deg (InitializeGraph2) or trim (KCoreStep1)
  unsigned long long* bits;
  PartCounters* c;
  uint32_t k;
  bool step;  // KCoreStep1
  unsigned active = 0;
  uint32_t* hot = nullptr;  // LDS counts of the pushes to ids below kHotIds
  __device__ __forceinline__ bool admit(uint32_t v) {
    if (!step) return true;
    if (flag[v] && deg[v] < k) {
      flag[v] = 0;
      ++active;
      return true;
    }
    return false;
  }
  __device__ __forceinline__ void operator()(bool valid, uint32_t, int32_t e) {
    if (!valid) return;
    const uint32_t v = static_cast<uint32_t>(col[e]);
    if (v < kHotIds) {  // counted in LDS, as the PageRank pushes
      atomicAdd(&hot[v], 1u);
      return;
    }
    atomicAdd(&target[v], 1u);
    set_bit(bits, v);
  }
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void finish() { flush_counts(c, active, 0ull); }
  // Workgroup-uniform.  Ids at or above n are never counted.
  __device__ __forceinline__ void flush_hot() {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHotIds; i += blockDim.x) {
      const uint32_t cnt = hot[i];
      if (cnt) {
        atomicAdd(&target[i], cnt);
        set_bit(bits, i);
      }
    }
  }
};

// Operators whose pushes to the hot ids are summed in LDS (flush_hot), with
// the type of their LDS sums.
template <class Op>
struct HotSums {
  static constexpr bool value = false;
  using T = uint32_t;
};
template <>
struct HotSums<PrPushOp> {
  static constexpr bool value = true;
  using T = float;
};
template <>
struct HotSums<KcoreOp> {
  static constexpr bool value = true;
  using T = uint32_t;
};

// KCoreStep2 (kcore_push.cpp:184-224): a live vertex loses
// its trim from its degree; trim is cleared everywhere in the range.
__global__ void kcore_step2_kernel(uint32_t* __restrict__ deg, uint32_t* __restrict__ trim,
                                   const uint32_t* __restrict__ flag, uint32_t begin, uint32_t end) {
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x) {
    const uint32_t t = trim[v];
    if (flag[v] && t > 0) deg[v] -= t;
    trim[v] = 0;
  }
}

// KCoreSanityCheck (kcore_push.cpp:321-365): live vertices in the range.
__global__ void __launch_bounds__(kBlock) kcore_check_kernel(const uint32_t* __restrict__ flag, uint32_t begin,
                                                             uint32_t end, PartCounters* c) {
  unsigned long long cnt = 0;
  for (uint32_t v = begin + blockIdx.x * blockDim.x + threadIdx.x; v < end; v += gridDim.x * blockDim.x)
    cnt += flag[v] ? 1 : 0;
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt) atomicAdd(&c->sum, cnt);
}

// ------------------------------------------------------- edge operators
template <class Op>
__global__ void __launch_bounds__(kBlock) range_expand_kernel(const int32_t* __restrict__ row, uint32_t begin,
                                                              uint32_t count, uint32_t hub, uint32_t* hubq,
                                                              PartCounters* c, unsigned cap, Op op) {
  if constexpr (HotSums<Op>::value) {
    __shared__ typename HotSums<Op>::T hot[kHotIds];
    for (uint32_t i = threadIdx.x; i < kHotIds; i += kBlock) hot[i] = 0;
    __syncthreads();
    op.hot = hot;
    expand_queue<kBlock, Op, true>(row, nullptr, count, hub, hubq, &c->hub_count, cap, &c->overflow, op, begin);
    op.flush_hot();
  } else {
    expand_queue<kBlock, Op, true>(row, nullptr, count, hub, hubq, &c->hub_count, cap, &c->overflow, op, begin);
  }
  op.finish();
}

template <class Op>
__global__ void __launch_bounds__(kBlock) range_hubs_kernel(const int32_t* __restrict__ row,
                                                            const uint32_t* __restrict__ hubq, PartCounters* c,
                                                            unsigned cap, const uint32_t* __restrict__ off, Op op) {
  if constexpr (HotSums<Op>::value) {
    __shared__ typename HotSums<Op>::T hot[kHotIds];
    for (uint32_t i = threadIdx.x; i < kHotIds; i += kBlock) hot[i] = 0;
    __syncthreads();
    op.hot = hot;
    expand_hubs<kBlock>(row, hubq, c->hub_count, cap, off, op);
    op.flush_hot();
  } else {
    expand_hubs<kBlock>(row, hubq, c->hub_count, cap, off, op);
  }
  op.finish();
}

unsigned grid_of(uint32_t count) {
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + kBlock - 1) / kBlock, kMaxGrid)));
}

void reset_counters(gx_part* p) {
  GX_HIP(hipMemsetAsync(p->counters.p, 0, sizeof(PartCounters), p->stream));
}

// Waits for the operator and copies its counters to the pinned host block.
const PartCounters& read_counters(gx_part* p) {
  GX_HIP(hipMemcpyAsync(p->host, p->counters.p, sizeof(PartCounters), hipMemcpyDeviceToHost, p->stream));
  GX_HIP(hipStreamSynchronize(p->stream));
  if (p->host->overflow) fail(GX_ERR_CAPACITY, "partition hub queue overflow");
  return *p->host;
}

// Range expansion of an edge operator: admitted vertices of [begin, end)
// expanded by workgroups, hubs spread over the grid.
template <class Op>
void run_edges(gx_part* p, uint32_t begin, uint32_t end, const Op& op) {
  const uint32_t count = end - begin;
  if (count == 0) return;
  PartCounters* c = p->counters.as<PartCounters>();
  const unsigned cap = static_cast<unsigned>(p->n);
  const int32_t* row = p->row.as<int32_t>();
  range_expand_kernel<Op><<<grid_of(count), kBlock, 0, p->stream>>>(row, begin, count, hub_degree(),
                                                                     p->hubq.as<uint32_t>(), c, cap, op);
  GX_LAUNCHED("range_expand_kernel");
  launch_hub_offsets<kBlock>(row, p->hubq.as<uint32_t>(), &c->hub_count, cap, p->hub_off.as<uint32_t>(), p->stream);
  GX_LAUNCHED("hub_offsets_kernel");
  range_hubs_kernel<Op><<<kMaxGrid, kBlock, 0, p->stream>>>(row, p->hubq.as<uint32_t>(), c, cap,
                                                           p->hub_off.as<uint32_t>(), op);
  GX_LAUNCHED("range_hubs_kernel");
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" gx_status gx_part_create(int device, gx_part** out) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  int count = 0;
  GX_HIP(hipGetDeviceCount(&count));
  GX_REQUIRE(device >= 0 && device < count, "device index out of range");
  GX_HIP(hipSetDevice(device));
  auto p = std::make_unique<gx_part>();
  p->device = device;
  GX_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  GX_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->host), sizeof(PartCounters), hipHostMallocDefault));
  p->counters.ensure(sizeof(PartCounters));
  *out = p.release();
  GX_API_END
}

extern "C" gx_status gx_part_free(gx_part* p) {
  GX_API_BEGIN
  if (p) {
    (void)hipStreamSynchronize(p->stream);
    delete p;
  }
  GX_API_END
}

extern "C" gx_status gx_part_load(gx_part* p, uint64_t n, uint64_t m, const uint32_t* row_start,
                                  const uint32_t* edge_dst, const uint32_t* edge_data, const uint32_t* node_gid,
                                  uint32_t num_owned, uint32_t begin_master, uint32_t num_nodes_with_edges) {
  GX_API_BEGIN
  GX_REQUIRE(p != nullptr, "partition handle is NULL");
  check_ptr(row_start, "row_start");
  GX_REQUIRE(n >= 1 && n < (uint64_t(1) << 31), "node count must be in [1, 2^31)");
  GX_REQUIRE(m < (uint64_t(1) << 31), "edge count must be < 2^31");
  if (m) check_ptr(edge_dst, "edge_dst");
  GX_REQUIRE(static_cast<uint64_t>(begin_master) + num_owned <= n, "master range outside the partition");
  GX_REQUIRE(num_nodes_with_edges <= n, "nodes-with-edges count exceeds the node count");
  // The kernels index by these values: validate before anything is uploaded.
  GX_REQUIRE(row_start[0] == 0 && row_start[n] == m, "row_start must start at 0 and end at m");
  for (uint64_t v = 0; v < n; ++v)
    GX_REQUIRE(row_start[v] <= row_start[v + 1], "row_start must be non-decreasing");
  for (uint64_t e = 0; e < m; ++e) GX_REQUIRE(edge_dst[e] < n, "edge destination out of range");
  GX_HIP(hipSetDevice(p->device));
  p->loaded = false;
  p->n = static_cast<int64_t>(n);
  p->m = static_cast<int64_t>(m);
  p->num_owned = num_owned;
  p->begin_master = begin_master;
  p->with_edges = num_nodes_with_edges;
  p->row.ensure((n + 1) * 4);
  p->col.ensure(std::max<uint64_t>(m, 1) * 4);
  GX_HIP(hipMemcpy(p->row.p, row_start, (n + 1) * 4, hipMemcpyHostToDevice));
  if (m) GX_HIP(hipMemcpy(p->col.p, edge_dst, m * 4, hipMemcpyHostToDevice));
  if (edge_data && m) {
    p->w.ensure(m * 4);
    GX_HIP(hipMemcpy(p->w.p, edge_data, m * 4, hipMemcpyHostToDevice));
  } else {
    p->w.release();
  }
  if (node_gid) {
    p->gid.ensure(n * 4);
    GX_HIP(hipMemcpy(p->gid.p, node_gid, n * 4, hipMemcpyHostToDevice));
  } else {
    p->gid.release();
  }
  for (int k = 0; k < kFields; ++k) {
    const uint64_t fb = k == GX_FIELD_BC_PATHS ? 8 : 4;  // path counts are FP64
    p->field[k].ensure(n * fb);
    GX_HIP(hipMemset(p->field[k].p, 0, n * fb));
    p->bits[k].ensure(p->words() * 8);
    GX_HIP(hipMemset(p->bits[k].p, 0, p->words() * 8));
  }
  p->hubq.ensure(n * 4);
  p->hub_off.ensure(hub_off_words(n) * 4);
  p->push.ensure(n * 4);
  p->bq.ensure(n * 4);
  p->loaded = true;
  GX_API_END
}

extern "C" gx_status gx_part_range(const gx_part* p, int which, uint32_t* begin, uint32_t* end) {
  GX_API_BEGIN
  check_part(p);
  check_ptr(begin, "begin");
  check_ptr(end, "end");
  switch (which) {
    case GX_RANGE_ALL: *begin = 0; *end = static_cast<uint32_t>(p->n); break;
    case GX_RANGE_MASTERS: *begin = p->begin_master; *end = p->begin_master + p->num_owned; break;
    case GX_RANGE_WITH_EDGES: *begin = 0; *end = p->with_edges; break;
    default: fail(GX_ERR_INVALID, "unknown range kind");
  }
  GX_API_END
}

extern "C" gx_status gx_part_get(const gx_part* p, int field, uint32_t lid, uint32_t* value) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_ptr(value, "value");
  check_range(p, lid, lid + 1);
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_HIP(hipMemcpy(value, p->f(field) + lid, 4, hipMemcpyDeviceToHost));
  GX_API_END
}

extern "C" gx_status gx_part_set(gx_part* p, int field, uint32_t lid, uint32_t value) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_range(p, lid, lid + 1);
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_HIP(hipMemcpy(p->f(field) + lid, &value, 4, hipMemcpyHostToDevice));
  GX_API_END
}

extern "C" gx_status gx_part_add(gx_part* p, int field, uint32_t lid, uint32_t value) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_range(p, lid, lid + 1);
  reset_counters(p);
  node_update_kernel<<<1, 64, 0, p->stream>>>(p->f(field), lid, value, 0, p->counters.as<PartCounters>());
  GX_LAUNCHED("node_update_kernel");
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_min(gx_part* p, int field, uint32_t lid, uint32_t value, int* changed) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_range(p, lid, lid + 1);
  reset_counters(p);
  node_update_kernel<<<1, 64, 0, p->stream>>>(p->f(field), lid, value, 1, p->counters.as<PartCounters>());
  GX_LAUNCHED("node_update_kernel");
  const PartCounters& c = read_counters(p);
  if (changed) *changed = static_cast<int>(c.changed);
  GX_API_END
}

extern "C" gx_status gx_part_fill(gx_part* p, int field, uint32_t begin, uint32_t end, uint32_t value) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_range(p, begin, end);
  if (end > begin) {
    fill_kernel_u32<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(field), begin, end, value);
    GX_LAUNCHED("fill_kernel_u32");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_read(const gx_part* p, int field, uint32_t begin, uint32_t end, uint32_t* host) {
  GX_API_BEGIN
  check_part(p);
  check_field(field);
  check_range(p, begin, end);
  if (end > begin) {
    check_ptr(host, "host");
    GX_HIP(hipStreamSynchronize(p->stream));
    GX_HIP(hipMemcpy(host, p->f(field) + begin, static_cast<size_t>(end - begin) * 4, hipMemcpyDeviceToHost));
  }
  GX_API_END
}

extern "C" gx_status gx_part_bitset_get(const gx_part* p, int field, uint64_t* words) {
  GX_API_BEGIN
  check_part(p);
  check_bitset_field(field);
  check_ptr(words, "words");
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_HIP(hipMemcpy(words, p->bits[field].p, p->words() * 8, hipMemcpyDeviceToHost));
  GX_API_END
}

extern "C" gx_status gx_part_bitset_reset(gx_part* p, int field, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_bitset_field(field);
  check_range(p, begin, end);
  if (end > begin) {
    const uint32_t nw = ((end - 1) >> 6) - (begin >> 6) + 1;
    bitset_clear_kernel<<<grid_of(nw), kBlock, 0, p->stream>>>(p->b(field), begin, end);
    GX_LAUNCHED("bitset_clear_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

static void push_edges(gx_part* p, uint32_t begin, uint32_t end, int kind, uint32_t priority, bool first) {
  const bool cc = kind == kPushCc;
  const int cur = cc ? GX_FIELD_COMP_CURRENT : GX_FIELD_DIST_CURRENT;
  const int old = cc ? GX_FIELD_COMP_OLD : GX_FIELD_DIST_OLD;
  if (kind == kPushSssp) GX_REQUIRE(p->w.p != nullptr, "SSSP push needs edge data (gx_part_load edge_data)");
  MinPushOp op{p->col.as<int32_t>(), p->w.as<uint32_t>(), p->f(cur), p->f(old), p->b(cur),
               p->counters.as<PartCounters>(), priority, kind, first};
  run_edges(p, begin, end, op);
}

static void dist_init(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t src_gid) {
  if (end > begin) {
    bfs_init_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_DIST_CURRENT), p->f(GX_FIELD_DIST_OLD), p->gid.p ? p->gid.as<uint32_t>() : nullptr, begin,
        end, infinity, src_gid);
    GX_LAUNCHED("bfs_init_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
}

static const PartCounters& dist_check(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity) {
  reset_counters(p);
  if (end > begin) {
    dist_check_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_DIST_CURRENT), begin, end,
                                                                      infinity, p->counters.as<PartCounters>());
    GX_LAUNCHED("dist_check_kernel");
  }
  return read_counters(p);
}

extern "C" gx_status gx_part_bfs_init(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity,
                                      uint64_t src_gid) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  dist_init(p, begin, end, infinity, src_gid);
  GX_API_END
}

extern "C" gx_status gx_part_bfs_first(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushBfs, 0, true);
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_bfs(gx_part* p, uint32_t begin, uint32_t end, uint32_t priority,
                                 uint32_t* active_vertices, uint32_t* work_items) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushBfs, priority, false);
  const PartCounters& c = read_counters(p);
  if (active_vertices) *active_vertices = c.active;
  if (work_items) *work_items = static_cast<uint32_t>(c.work);
  GX_API_END
}

extern "C" gx_status gx_part_bfs_check(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity,
                                       uint64_t* sum, uint32_t* max_dist) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  const PartCounters& c = dist_check(p, begin, end, infinity);
  if (sum) *sum = c.sum;
  if (max_dist) *max_dist = c.max;
  GX_API_END
}

extern "C" gx_status gx_part_sssp_init(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity,
                                       uint64_t src_gid) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  dist_init(p, begin, end, infinity, src_gid);
  GX_API_END
}

extern "C" gx_status gx_part_sssp_first(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushSssp, 0, true);
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_sssp(gx_part* p, uint32_t begin, uint32_t end, uint32_t priority,
                                  uint32_t* active_vertices, uint32_t* work_edges) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushSssp, priority, false);
  const PartCounters& c = read_counters(p);
  if (active_vertices) *active_vertices = c.active;
  if (work_edges) *work_edges = static_cast<uint32_t>(c.work);
  GX_API_END
}

extern "C" gx_status gx_part_sssp_check(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity,
                                        uint64_t* sum, uint64_t* dist_sum, uint32_t* max_dist) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  const PartCounters& c = dist_check(p, begin, end, infinity);
  if (sum) *sum = c.sum;
  if (dist_sum) *dist_sum = c.dsum;
  if (max_dist) *max_dist = c.max;
  GX_API_END
}

extern "C" gx_status gx_part_cc_init(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    cc_init_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_COMP_CURRENT), p->f(GX_FIELD_COMP_OLD), p->gid.p ? p->gid.as<uint32_t>() : nullptr, begin,
        end);
    GX_LAUNCHED("cc_init_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_cc_first(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushCc, 0, true);
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_cc(gx_part* p, uint32_t begin, uint32_t end, uint32_t* active_vertices) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  push_edges(p, begin, end, kPushCc, 0, false);
  const PartCounters& c = read_counters(p);
  if (active_vertices) *active_vertices = c.active;
  GX_API_END
}

extern "C" gx_status gx_part_cc_check(gx_part* p, uint32_t begin, uint32_t end, uint64_t* components) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  if (end > begin) {
    cc_check_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_COMP_CURRENT), p->gid.p ? p->gid.as<uint32_t>() : nullptr, begin, end,
        p->counters.as<PartCounters>());
    GX_LAUNCHED("cc_check_kernel");
  }
  const PartCounters& c = read_counters(p);
  if (components) *components = c.sum;
  GX_API_END
}

extern "C" gx_status gx_part_pr_reset(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    pr_reset_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_PR_VALUE), p->f(GX_FIELD_PR_NOUT), p->f(GX_FIELD_PR_RESIDUAL), p->f(GX_FIELD_PR_DELTA), begin,
        end);
    GX_LAUNCHED("pr_reset_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_pr_init(gx_part* p, uint32_t begin, uint32_t end, float alpha) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    pr_init_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->row.as<int32_t>(), p->f(GX_FIELD_PR_NOUT), p->f(GX_FIELD_PR_RESIDUAL), p->b(GX_FIELD_PR_NOUT), alpha,
        begin, end);
    GX_LAUNCHED("pr_init_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_pr_delta(gx_part* p, uint32_t begin, uint32_t end, float alpha, float tolerance) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    pr_delta_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_PR_VALUE), p->f(GX_FIELD_PR_NOUT), p->f(GX_FIELD_PR_RESIDUAL), p->f(GX_FIELD_PR_DELTA), alpha,
        tolerance, begin, end);
    GX_LAUNCHED("pr_delta_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_pr(gx_part* p, uint32_t begin, uint32_t end, uint32_t* active_vertices) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  PrPushOp op{p->col.as<int32_t>(), p->f(GX_FIELD_PR_DELTA), p->push.as<float>(),
              reinterpret_cast<float*>(p->f(GX_FIELD_PR_RESIDUAL)), p->b(GX_FIELD_PR_RESIDUAL),
              p->counters.as<PartCounters>()};
  run_edges(p, begin, end, op);
  const PartCounters& c = read_counters(p);
  if (active_vertices) *active_vertices = c.active;
  GX_API_END
}

extern "C" gx_status gx_part_pr_check(gx_part* p, uint32_t begin, uint32_t end, float tolerance,
                                      gx_pr_sanity* out) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  check_ptr(out, "out");
  // Minima start at +inf (bit pattern); staged through the pinned block,
  // which the stream reads before read_counters writes it again.
  *p->host = PartCounters{};
  p->host->min_value = p->host->min_res = 0x7f800000u;
  GX_HIP(hipMemcpyAsync(p->counters.p, p->host, sizeof(PartCounters), hipMemcpyHostToDevice, p->stream));
  if (end > begin) {
    pr_check_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_PR_VALUE), p->f(GX_FIELD_PR_RESIDUAL), tolerance, begin, end, p->counters.as<PartCounters>());
    GX_LAUNCHED("pr_check_kernel");
  }
  const PartCounters& c = read_counters(p);
  auto f = [](uint32_t b) {
    float x;
    std::memcpy(&x, &b, 4);
    return x;
  };
  out->max_value = f(c.max_value);
  out->min_value = f(c.min_value);
  out->max_residual = f(c.max_res);
  out->min_residual = f(c.min_res);
  out->sum_value = c.sum_value;
  out->sum_residual = c.sum_res;
  out->residual_over_tolerance = c.over_tol;
  GX_API_END
}

static double* paths(gx_part* p) { return p->field[GX_FIELD_BC_PATHS].as<double>(); }

extern "C" gx_status gx_part_bc_init(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    bc_init_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_BC_VALUE),
                                                                    p->f(GX_FIELD_BC_DEPENDENCY), paths(p), begin, end);
    GX_LAUNCHED("bc_init_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_bc_iter(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity,
                                     uint64_t src_gid) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    bc_iter_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_BC_LENGTH), paths(p), p->f(GX_FIELD_BC_DEPENDENCY), p->gid.p ? p->gid.as<uint32_t>() : nullptr,
        begin, end, infinity, src_gid);
    GX_LAUNCHED("bc_iter_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_bc_forward(gx_part* p, uint32_t begin, uint32_t end, uint32_t round, uint32_t* dga) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  BcForwardOp op{p->col.as<int32_t>(), p->f(GX_FIELD_BC_LENGTH), paths(p), p->b(GX_FIELD_BC_LENGTH),
                 p->b(GX_FIELD_BC_PATHS), p->counters.as<PartCounters>(), round};
  run_edges(p, begin, end, op);
  const PartCounters& c = read_counters(p);
  if (dga) *dga = c.active;
  GX_API_END
}

extern "C" gx_status gx_part_bc_middle(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    bc_middle_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_BC_LENGTH),
                                                                      p->b(GX_FIELD_BC_PATHS), begin, end, infinity);
    GX_LAUNCHED("bc_middle_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_bc_backward(gx_part* p, uint32_t begin, uint32_t end, uint32_t round) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  if (end > begin) {
    PartCounters* c = p->counters.as<PartCounters>();
    bc_level_queue_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_BC_LENGTH), begin, end,
                                                                         round, p->bq.as<uint32_t>(), c);
    GX_LAUNCHED("bc_level_queue_kernel");
    bc_backward_kernel<<<kMaxGrid, kBlock, 0, p->stream>>>(
        p->row.as<int32_t>(), p->col.as<int32_t>(), p->f(GX_FIELD_BC_LENGTH), paths(p), p->f(GX_FIELD_BC_DEPENDENCY),
        p->b(GX_FIELD_BC_DEPENDENCY), p->bq.as<uint32_t>(), c, round);
    GX_LAUNCHED("bc_backward_kernel");
  }
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_bc_accumulate(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    bc_accumulate_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_BC_VALUE),
                                                                          p->f(GX_FIELD_BC_DEPENDENCY), begin, end);
    GX_LAUNCHED("bc_accumulate_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_bc_check(gx_part* p, uint32_t begin, uint32_t end, float* sum, float* max_bc,
                                      float* min_bc) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  *p->host = PartCounters{};
  p->host->min_value = 0x7f800000u;
  GX_HIP(hipMemcpyAsync(p->counters.p, p->host, sizeof(PartCounters), hipMemcpyHostToDevice, p->stream));
  if (end > begin) {
    bc_check_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_BC_VALUE), begin, end,
                                                                     p->counters.as<PartCounters>());
    GX_LAUNCHED("bc_check_kernel");
  }
  const PartCounters& c = read_counters(p);
  float mx, mn;
  std::memcpy(&mx, &c.max_value, 4);
  std::memcpy(&mn, &c.min_value, 4);
  if (sum) *sum = static_cast<float>(c.sum_value);
  if (max_bc) *max_bc = mx;
  if (min_bc) *min_bc = mn;
  GX_API_END
}

extern "C" gx_status gx_part_read_paths(const gx_part* p, uint32_t begin, uint32_t end, double* host) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    check_ptr(host, "host");
    GX_HIP(hipStreamSynchronize(p->stream));
    GX_HIP(hipMemcpy(host, p->field[GX_FIELD_BC_PATHS].as<double>() + begin, static_cast<size_t>(end - begin) * 8,
                     hipMemcpyDeviceToHost));
  }
  GX_API_END
}

extern "C" gx_status gx_part_kcore_init1(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    kcore_init1_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_CURRENT_DEGREE), p->f(GX_FIELD_TRIM), p->f(GX_FIELD_FLAG), begin, end);
    GX_LAUNCHED("kcore_init1_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

static void kcore_edges(gx_part* p, uint32_t begin, uint32_t end, uint32_t k, bool step) {
  const int target = step ? GX_FIELD_TRIM : GX_FIELD_CURRENT_DEGREE;
  KcoreOp op{p->col.as<int32_t>(), p->f(GX_FIELD_CURRENT_DEGREE), p->f(GX_FIELD_TRIM), p->f(GX_FIELD_FLAG),
             p->f(target), p->b(target), p->counters.as<PartCounters>(), k, step};
  run_edges(p, begin, end, op);
}

extern "C" gx_status gx_part_kcore_init2(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  kcore_edges(p, begin, end, 0, false);
  read_counters(p);
  GX_API_END
}

extern "C" gx_status gx_part_kcore_step1(gx_part* p, uint32_t begin, uint32_t end, uint32_t k,
                                         uint32_t* active_vertices) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  kcore_edges(p, begin, end, k, true);
  const PartCounters& c = read_counters(p);
  if (active_vertices) *active_vertices = c.active;
  GX_API_END
}

extern "C" gx_status gx_part_kcore_step2(gx_part* p, uint32_t begin, uint32_t end) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  if (end > begin) {
    kcore_step2_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(
        p->f(GX_FIELD_CURRENT_DEGREE), p->f(GX_FIELD_TRIM), p->f(GX_FIELD_FLAG), begin, end);
    GX_LAUNCHED("kcore_step2_kernel");
  }
  GX_HIP(hipStreamSynchronize(p->stream));
  GX_API_END
}

extern "C" gx_status gx_part_kcore_check(gx_part* p, uint32_t begin, uint32_t end, uint64_t* alive) {
  GX_API_BEGIN
  check_part(p);
  check_range(p, begin, end);
  reset_counters(p);
  if (end > begin) {
    kcore_check_kernel<<<grid_of(end - begin), kBlock, 0, p->stream>>>(p->f(GX_FIELD_FLAG), begin, end,
                                                                       p->counters.as<PartCounters>());
    GX_LAUNCHED("kcore_check_kernel");
  }
  const PartCounters& c = read_counters(p);
  if (alive) *alive = c.sum;
  GX_API_END
}
