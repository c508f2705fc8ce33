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
// Design (gfx950): the sampling scheme of the reference's Afforest variant
// (ConnectedComponents.cpp:654-759) on a lock-free GPU union-find:
//  1. link every vertex with its first two neighbours (two vertex-parallel
//     passes), then compress;
//  2. estimate the largest component from 1024 hashed samples of roots;
//  3. hook the remaining edges of every vertex outside that component (load
//     balanced with frontier.hpp over a compacted vertex list; hubs spread
//     over the grid).  Edges of vertices inside it need no hooking: the graph
//     is symmetric, so their other endpoint hooks the edge if it matters;
//  4. compress.
// find() uses path halving with relaxed agent-scope atomics; linking is one
// atomicCAS of the higher root onto the lower one, so roots stay component
// minima and the labels are the canonical min labels.
#include <algorithm>
#include <mutex>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;

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

// Sampling pass r: link v with its r-th neighbour.
__global__ void cc_sample_kernel(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                                 uint32_t* __restrict__ parent, int64_t n, int r) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t b = row[v];
    if (row[v + 1] - b > r) unite(parent, static_cast<uint32_t>(v), static_cast<uint32_t>(col[b + r]));
  }
}

// Roots of 1024 hashed sample vertices; the most frequent one is taken as the
// largest component (single workgroup, counts by comparison in LDS).
__global__ void __launch_bounds__(256) cc_largest_kernel(uint32_t* __restrict__ parent, int64_t n,
                                                         uint32_t* __restrict__ out) {
  constexpr int kSamples = 1024;
  __shared__ uint32_t roots[kSamples];
  __shared__ unsigned best_cnt[256];
  __shared__ uint32_t best_root[256];
  for (int i = threadIdx.x; i < kSamples; i += 256) {
    const uint64_t v = (mix64(0x43434343u + static_cast<uint64_t>(i)) >> 11) % static_cast<uint64_t>(n);
    roots[i] = find_root(parent, static_cast<uint32_t>(v));
  }
  __syncthreads();
  unsigned bc = 0;
  uint32_t br = 0;
  for (int i = threadIdx.x; i < kSamples; i += 256) {
    unsigned c = 0;
    for (int j = 0; j < kSamples; ++j) c += roots[j] == roots[i];
    if (c > bc || (c == bc && roots[i] < br)) {
      bc = c;
      br = roots[i];
    }
  }
  best_cnt[threadIdx.x] = bc;
  best_root[threadIdx.x] = br;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < 256; ++t)
      if (best_cnt[t] > bc || (best_cnt[t] == bc && best_root[t] < br)) {
        bc = best_cnt[t];
        br = best_root[t];
      }
    *out = br;
  }
}

constexpr int kQCap = 2048;

// Vertices outside the largest component that still have unhooked edges.
__global__ void __launch_bounds__(kBlock) cc_rest_list_kernel(const int32_t* __restrict__ row,
                                                              uint32_t* __restrict__ parent, int64_t n,
                                                              const uint32_t* __restrict__ largest,
                                                              uint32_t* __restrict__ q, unsigned* count,
                                                              unsigned cap, unsigned* overflow) {
  __shared__ LdsQueue<kQCap> sq;
  lq_init(sq);
  const uint32_t c = *largest;
  for (int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x); b < n;
       b += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t v = b + threadIdx.x;
    bool keep = false;
    if (v < n && row[v + 1] - row[v] > 2) keep = find_root(parent, static_cast<uint32_t>(v)) != c;
    lq_push(sq, keep, static_cast<uint32_t>(v));
    lq_flush(sq, q, count, cap, overflow, false);
  }
  lq_flush(sq, q, count, cap, overflow, true);
}

struct HookOp {
  const int32_t* row;
  const int32_t* col;
  uint32_t* parent;
  __device__ __forceinline__ void after_step() {}
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    if (!valid || e - row[u] < 2) return;  // the first two edges were sampled
    unite(parent, u, static_cast<uint32_t>(col[e]));
  }
};

__global__ void __launch_bounds__(kBlock) cc_hook_kernel(const int32_t* __restrict__ row, HookOp op,
                                                         const uint32_t* __restrict__ q, const unsigned* count,
                                                         uint32_t* __restrict__ hubq, unsigned* hub_count,
                                                         unsigned cap, unsigned* overflow, uint32_t hub) {
  expand_queue<kBlock>(row, q, *count, hub, hubq, hub_count, cap, overflow, op);
}

__global__ void __launch_bounds__(kBlock) cc_hubs_kernel(const int32_t* __restrict__ row, HookOp op,
                                                         const uint32_t* __restrict__ hubq,
                                                         const unsigned* hub_count, unsigned cap,
                                                         const uint32_t* __restrict__ hub_off) {
  expand_hubs<kBlock>(row, hubq, *hub_count, cap, hub_off, op);
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
  uint32_t* rest = g->scratch<uint32_t>(WS_Q1, cap);
  uint32_t* hubq = g->scratch<uint32_t>(WS_Q2, cap);
  uint32_t* hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(cap));
  unsigned* ctr = g->scratch<unsigned>(WS_COUNTERS, 4);  // [0] hubs, [1] overflow, [2] rest, [3] largest
  RegionTimer region(s);
  KernelTimer ktimer(s);
  const unsigned vgrid = grid_for(n, kBlock, 4096);
  if (n > 0) {
    cc_zero_kernel<<<1, 64, 0, s>>>(ctr);
    cc_init_kernel<<<vgrid, kBlock, 0, s>>>(parent, n);
    GX_LAUNCHED("cc_init_kernel");
    ktimer.begin();
    for (int r = 0; r < 2; ++r) cc_sample_kernel<<<vgrid, kBlock, 0, s>>>(g->row, g->col, parent, n, r);
    cc_compress_kernel<<<vgrid, kBlock, 0, s>>>(parent, label, n);
    cc_largest_kernel<<<1, 256, 0, s>>>(parent, n, &ctr[3]);
    cc_rest_list_kernel<<<vgrid, kBlock, 0, s>>>(g->row, parent, n, &ctr[3], rest, &ctr[2], cap, &ctr[1]);
    HookOp op{g->row, g->col, parent};
    cc_hook_kernel<<<256 * 8, kBlock, 0, s>>>(g->row, op, rest, &ctr[2], hubq, &ctr[0], cap, &ctr[1],
                                                 hub_degree());
    GX_LAUNCHED("cc_hook_kernel");
    launch_hub_offsets<kBlock>(g->row, hubq, &ctr[0], cap, hub_off, s);
    cc_hubs_kernel<<<256 * 8, kBlock, 0, s>>>(g->row, op, hubq, &ctr[0], cap, hub_off);
    ktimer.end();
    GX_LAUNCHED("cc_hubs_kernel");
    cc_compress_kernel<<<vgrid, kBlock, 0, s>>>(parent, label, n);
    GX_LAUNCHED("cc_compress_kernel");
    unsigned hc[2];
    read_bytes(g, hc, ctr, sizeof(hc));
    if (hc[1]) fail(GX_ERR_CAPACITY, "CC vertex list overflow");
  }
  double ms = region.stop();
  if (stp) {
    std::memset(stp, 0, sizeof(*stp));
    stp->iterations = 1;
    stp->launches = n > 0 ? 10 : 0;
    stp->edges_visited = static_cast<uint64_t>(g->m);
    stp->items = static_cast<uint64_t>(n);
    stp->elapsed_ms = ms;
    stp->main_kernel_ms = ktimer.total_ms();
    stp->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}

// ------------------------------------------------------------ no allocation
// gx_cc_noalloc runs the kernels of gx_cc with their scratch (union-find
// parents, the vertex list, the hub list and its offsets, four counters) in
// one file-scope device array instead of the graph's workspace, so a call
// allocates nothing.  Each device holds its own copy of the array from module
// load until the process ends; its address on the current device is resolved
// once per device (gx_init, or the first call on that device) and kept.
namespace {

constexpr int64_t kNoAllocMaxVertices = int64_t(1) << 22;
constexpr int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }
// Byte offsets of the scratch parts for n vertices: counters, parents, vertex
// list, hub list, hub offsets.
struct NoAllocParts {
  int64_t parent, rest, hubq, hub_off, end;
};
constexpr NoAllocParts noalloc_parts(int64_t n) {
  const int64_t w = align256(4 * n);
  return NoAllocParts{256, 256 + w, 256 + 2 * w, 256 + 3 * w, 256 + 3 * w + align256(4 * hub_off_words(n))};
}
constexpr int64_t kNoAllocScratchBytes = noalloc_parts(kNoAllocMaxVertices).end;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace gx {
unsigned char* noalloc_scratch() {
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  GX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GX_REQUIRE(p != nullptr, "no device address for the no-allocation scratch");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace gx

extern "C" gx_status gx_cc_noalloc(gx_graph* g, uint32_t* label) {
  GX_API_BEGIN
  check_graph(g);
  check_ptr(label, "label");
  GX_REQUIRE(g->flags & GX_GRAPH_SYMMETRIC, "connected components needs a symmetric graph");
  const int64_t n = g->n;
  if (n > kNoAllocMaxVertices)
    fail(GX_ERR_CAPACITY, "gx_cc_noalloc: the graph has " + std::to_string(n) +
                              " vertices; its static scratch holds at most " + std::to_string(kNoAllocMaxVertices));
  if (n == 0) return GX_OK;
  hipStream_t s = g->stream;
  unsigned char* base = noalloc_scratch();  // this device's address of `scratch`
  const NoAllocParts P = noalloc_parts(n);
  unsigned* ctr = reinterpret_cast<unsigned*>(base);  // [0] hubs, [1] overflow, [2] rest, [3] largest
  uint32_t* parent = reinterpret_cast<uint32_t*>(base + P.parent);
  uint32_t* rest = reinterpret_cast<uint32_t*>(base + P.rest);
  uint32_t* hubq = reinterpret_cast<uint32_t*>(base + P.hubq);
  uint32_t* hub_off = reinterpret_cast<uint32_t*>(base + P.hub_off);
  const unsigned cap = static_cast<unsigned>(n);  // lists hold at most n vertices: no overflow
  const unsigned vgrid = grid_for(n, kBlock, 4096);
  cc_zero_kernel<<<1, 64, 0, s>>>(ctr);
  cc_init_kernel<<<vgrid, kBlock, 0, s>>>(parent, n);
  GX_LAUNCHED("cc_init_kernel");
  for (int r = 0; r < 2; ++r) cc_sample_kernel<<<vgrid, kBlock, 0, s>>>(g->row, g->col, parent, n, r);
  cc_compress_kernel<<<vgrid, kBlock, 0, s>>>(parent, label, n);
  cc_largest_kernel<<<1, 256, 0, s>>>(parent, n, &ctr[3]);
  cc_rest_list_kernel<<<vgrid, kBlock, 0, s>>>(g->row, parent, n, &ctr[3], rest, &ctr[2], cap, &ctr[1]);
  HookOp op{g->row, g->col, parent};
  cc_hook_kernel<<<256 * 8, kBlock, 0, s>>>(g->row, op, rest, &ctr[2], hubq, &ctr[0], cap, &ctr[1], hub_degree());
  GX_LAUNCHED("cc_hook_kernel");
  launch_hub_offsets<kBlock>(g->row, hubq, &ctr[0], cap, hub_off, s);
  cc_hubs_kernel<<<256 * 8, kBlock, 0, s>>>(g->row, op, hubq, &ctr[0], cap, hub_off);
  GX_LAUNCHED("cc_hubs_kernel");
  cc_compress_kernel<<<vgrid, kBlock, 0, s>>>(parent, label, n);
  GX_LAUNCHED("cc_compress_kernel");
  GX_HIP(hipStreamSynchronize(s));
  GX_API_END
}
