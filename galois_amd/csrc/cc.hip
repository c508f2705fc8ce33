// Connected components (symmetric graphs) with a lock-free union-find.
//
// Reference semantics: the union-find of libgalois/include/galois/UnionFind.h
// (merge :108-122 links the root with the higher address into the lower
// one; compress :66-82) driven over every edge with src < dst by
// lonestar/analytics/cpu/connected-components/ConnectedComponents.cpp
// (EdgeTiledAsync :523-613, Async :341-383).  Because a root is only ever
// linked below a smaller root, each tree's root is the minimum vertex id of
// its set, and after full compression label[v] = min id of v's component —
// the canonical labelling that LabelProp (:183-230) also produces.  The GPU
// app's hook/jump scheme (lonestar/analytics/gpu/connected-components/
// cc.cu:244-357) leaves non-minimum roots; this kernel set does not.
//
// Design (gfx950): one hooking pass over all edges, load balanced with the
// workgroup-local expansion of frontier.hpp over the identity frontier (hubs
// spread over the grid); find() uses path halving with relaxed agent-scope
// atomics; linking is one atomicCAS on the higher root.  A final sweep writes
// the fully compressed root of every vertex.
#include <algorithm>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kHub = 4096;

__device__ __forceinline__ uint32_t ld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving: every visited node is re-pointed to its
// grandparent.  Only non-roots are rewritten and always to an ancestor, so the
// writes are benign under concurrent linking.
__device__ __forceinline__ uint32_t find_root(uint32_t* parent, uint32_t x) {
  uint32_t p = ld(&parent[x]);
  while (p != x) {
    const uint32_t gp = ld(&parent[p]);
    if (gp != p) st(&parent[x], gp);
    x = gp;
    p = ld(&parent[x]);
  }
  return x;
}

__device__ __forceinline__ void unite(uint32_t* parent, uint32_t a, uint32_t b) {
  for (;;) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // a is the higher root: link it below b.
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

__global__ void cc_init_kernel(uint32_t* __restrict__ parent, int64_t n) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) parent[v] = static_cast<uint32_t>(v);
}

struct HookOp {
  const int32_t* col;
  uint32_t* parent;
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid) return;
    const uint32_t v = static_cast<uint32_t>(col[e]);
    if (v < u) unite(parent, u, v);
  }
};

__global__ void __launch_bounds__(kBlock) cc_hook_kernel(const int32_t* __restrict__ row, HookOp op, unsigned n,
                                                         uint32_t* __restrict__ hubq, unsigned* hub_count,
                                                         unsigned* overflow) {
  expand_queue<kBlock>(row, nullptr, n, kHub, hubq, hub_count, n, overflow, op);
}

__global__ void __launch_bounds__(kBlock) cc_hubs_kernel(const int32_t* __restrict__ row, HookOp op,
                                                         const uint32_t* __restrict__ hubq,
                                                         const unsigned* hub_count) {
  expand_hubs<kBlock>(row, hubq, *hub_count, op);
}

__global__ void cc_compress_kernel(uint32_t* __restrict__ parent, uint32_t* __restrict__ label, int64_t n) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) label[v] = find_root(parent, static_cast<uint32_t>(v));
}

__global__ void cc_zero_kernel(unsigned* c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

}  // namespace

extern "C" gx_status gx_cc(gx_graph* g, uint32_t* label, gx_stats* stp) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(label, "label");
  GX_REQUIRE(g->flags & GX_GRAPH_SYMMETRIC, "connected components needs a symmetric graph");
  const int64_t n = g->n;
  hipStream_t s = g->stream;
  const unsigned cap = static_cast<unsigned>(std::max<int64_t>(n, 1));
  uint32_t* parent = g->scratch<uint32_t>(WS_V0, cap);
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, cap);
  unsigned* ctr = g->scratch<unsigned>(WS_COUNTERS, 4);  // [0] hubs, [1] overflow
  RegionTimer region(s);
  KernelTimer ktimer(s);
  if (n > 0) {
    cc_zero_kernel<<<1, 64, 0, s>>>(ctr);
    cc_init_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(parent, n);
    GX_LAUNCHED("cc_init_kernel");
    HookOp op{g->col, parent};
    ktimer.begin();
    cc_hook_kernel<<<grid_for(n, kBlock, 256 * 16), kBlock, 0, s>>>(g->row, op, static_cast<unsigned>(n), hubq,
                                                                      &ctr[0], &ctr[1]);
    GX_LAUNCHED("cc_hook_kernel");
    cc_hubs_kernel<<<1024, kBlock, 0, s>>>(g->row, op, hubq, &ctr[0]);
    ktimer.end();
    GX_LAUNCHED("cc_hubs_kernel");
    cc_compress_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(parent, label, n);
    GX_LAUNCHED("cc_compress_kernel");
  }
  double ms = region.stop();
  if (stp) {
    std::memset(stp, 0, sizeof(*stp));
    stp->iterations = 1;
    stp->launches = n > 0 ? 5 : 0;
    stp->edges_visited = static_cast<uint64_t>(g->m);
    stp->items = static_cast<uint64_t>(n);
    stp->elapsed_ms = ms;
    stp->main_kernel_ms = ktimer.total_ms();
    stp->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
