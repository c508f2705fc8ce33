// Internal host-side definitions shared by every translation unit of
// libgalois_amd.so: the device graph object, error plumbing across the C ABI,
// the per-graph scratch cache and the HIP-event profiling helper.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/galois_amd.h"

namespace gx {

// ---------------------------------------------------------------- errors
// Internal code throws gx::Error; every extern "C" entry point catches it at
// the boundary (GX_API_BEGIN/END), records the message and returns the code.
struct Error : std::runtime_error {
  gx_status code;
  Error(gx_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(gx_status c, const std::string& msg) { throw Error(c, msg); }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    fail(e == hipErrorOutOfMemory ? GX_ERR_ALLOC : GX_ERR_HIP,
         std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")");
  }
}
#define GX_HIP(expr) ::gx::hip_check((expr), #expr, __FILE__, __LINE__)
// Launch-error check after every kernel launch.
#define GX_LAUNCHED(name) ::gx::hip_check(hipGetLastError(), name, __FILE__, __LINE__)

#define GX_REQUIRE(cond, msg)                        \
  do {                                               \
    if (!(cond)) ::gx::fail(GX_ERR_INVALID, (msg)); \
  } while (0)

#define GX_API_BEGIN try {
#define GX_API_END                                     \
  }                                                    \
  catch (const ::gx::Error& e) {                       \
    ::gx::set_last_error(e.what());                    \
    return e.code;                                     \
  }                                                    \
  catch (const std::bad_alloc&) {                      \
    ::gx::set_last_error("host allocation failed");    \
    return GX_ERR_ALLOC;                               \
  }                                                    \
  catch (const std::exception& e) {                    \
    ::gx::set_last_error(e.what());                    \
    return GX_ERR_INVALID;                             \
  }                                                    \
  return GX_OK;

// ------------------------------------------------------------ device memory
// Owning device buffer (hipMalloc/hipFree), move-only.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grow to at least `need` bytes (contents are not preserved).
  void ensure(size_t need) {
    if (need <= bytes) return;
    release();
    size_t b = need < 256 ? 256 : need;
    GX_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Scratch slots a graph keeps between calls so that repeated runs allocate
// nothing.  Each algorithm uses its own slots.
enum Slot : int {
  WS_Q0 = 0, WS_Q1, WS_Q2, WS_Q3, WS_COUNTERS, WS_BITMAP0, WS_BITMAP1,
  WS_V0, WS_V1, WS_V2, WS_V3, WS_V4, WS_V5, WS_REDUCE, WS_BC_LEVEL, WS_BC_SIGMA,
  WS_BC_DELTA, WS_BC_MASK0, WS_BC_MASK1, WS_BC_MASK2, WS_HOST_PINNED, WS_HUBOFF,
  WS_NUM_SLOTS
};

// The transpose's column array is allocated with this many trailing entries
// (filled with vertex 0) so vectorised edge loads may run past m.
constexpr int64_t kEdgePad = 4096;

// Tables of the pull PageRank sweep over the transpose, built once per graph.
struct PullPlan {
  bool ready = false;
  int64_t ntiles = 0;
  DevBuf buf;
};

// The pull graph of PageRank in degree-sorted vertex order: new id = rank of
// the vertex by (out-degree descending, id ascending).  Sources that are
// gathered most often get the lowest ids, so the hot part of the gathered
// vector is a few contiguous megabytes.  trow/tcol are the transpose in new
// ids (tcol padded by kEdgePad), outdeg is in new order.
// Edge layout of the propagation-blocked PageRank sweep (pr_blocking.hip),
// built once per graph from the CSR in degree-order ids.  Edges are sorted
// into runs by (source chunk, destination bin), chunk-major, and by (source,
// destination) inside a run; a group is the edges of one source in a run.
// dst16 is the in-bin destination of every edge (bit 15: first edge of a
// group), stored with the runs in bin-major order (the runs of one bin from
// every chunk are adjacent); gsrc16 is the in-chunk source of every group,
// in run-major order, and vals the per-group
// contributions of the current iteration.  pieces/jobs/tasks/multi are the
// work tables of the two phases.
struct PbLayout {
  bool ready = false;
  int64_t nbins = 0, ngroups = 0, npieces = 0, ntasks = 0, nmulti = 0;
  int64_t njobs[3] = {0, 0, 0};  // long, mid, short jobs
  int64_t task_entries = 0;      // slice size (edges) the tasks were cut with
  int64_t nonzero = 0;           // vertices with out-edges: degree-order ids [0, nonzero)
  DevBuf gsrc16, dst16, vals, pieces, jobs, tasks, multi, gacc;
};

struct DegreeOrder {
  bool ready = false;
  DevBuf new_of_old, old_of_new, trow, tcol, outdeg;
  // Hot/cold split of the transpose rows by source id (< hot / >= hot):
  // two CSRs over the same rows (PageRank pull).
  int64_t hot = -1, hot_edges = 0;
  DevBuf rowA, colA, rowB, colB;
  PullPlan planA, planB;
  PbLayout pb;
};

// Degree-ordered copy of a weighted CSR for the frontier kernels: vertices
// renumbered by (degree descending, id ascending), out-edges sorted by new
// target id, weights carried along.  The most often relaxed targets (the
// hubs and their neighbourhoods) then occupy a few contiguous megabytes of
// a per-vertex array, which stay in L2 instead of costing a line fetch per
// access.  Built on demand, kept with the graph.
struct Relabeled {
  bool ready = false;
  DevBuf new_of_old, old_of_new, row, col, w;
};

// A fixed kernel sequence captured once into an executable HIP graph and
// replayed; re-captured when `key` (identity of the baked kernel arguments)
// changes.  Used for the device-driven level/bucket loops.
struct CapturedBatch {
  hipGraphExec_t exec = nullptr;
  uint64_t key = 0;
  CapturedBatch() = default;
  CapturedBatch(const CapturedBatch&) = delete;
  CapturedBatch& operator=(const CapturedBatch&) = delete;
  ~CapturedBatch() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key = 0;
  }
};

inline uint64_t mix_key(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

}  // namespace gx

// The opaque graph handle of the C ABI.
struct gx_graph {
  int64_t n = 0, m = 0;
  uint32_t flags = 0;
  // CSR (out-edges)
  int32_t* row = nullptr;
  int32_t* col = nullptr;
  void* w = nullptr;  // uint32 or float weights (m) or null
  bool owns_csr = true;
  gx::DevBuf row_buf, col_buf, w_buf;
  // CSC (in-edges), built on demand; tperm maps CSC slot -> CSR edge id.
  gx::DevBuf trow_buf, tcol_buf, tperm_buf;
  bool has_transpose = false;
  // Derived per-vertex data, built on demand.
  gx::DevBuf outdeg_buf;  // uint32 out-degree (for transpose-based PR)
  bool has_outdeg = false;
  int64_t max_deg = -1;  // largest out-degree, computed on demand (gx::max_degree)
  gx::PullPlan pull_plan;            // degree-ordered transpose (PageRank pull)
  gx::PullPlan csc_plan, csr_plan;   // plain transpose / CSR (weighted PageRank and adjoint)
  gx::DegreeOrder degree_order;
  gx::Relabeled relab;
  std::shared_ptr<void> bc_cache;  // batched-BC tables (bc.hip)
  hipStream_t stream = nullptr;
  hipStream_t capture_stream = nullptr;  // non-null stream for HIP graph capture
  gx::CapturedBatch sssp_batch, sssp_batched_batch, bfs_batch, bfs_head, kcore_batch, kcore_head;
  gx::DevBuf ws[gx::WS_NUM_SLOTS];
  void* pinned = nullptr;  // small pinned host block for counters
  size_t pinned_bytes = 0;
  ~gx_graph();

  int32_t* trow() const { return trow_buf.as<int32_t>(); }
  int32_t* tcol() const { return tcol_buf.as<int32_t>(); }
  int32_t* tperm() const { return tperm_buf.as<int32_t>(); }
  template <class T> T* scratch(int slot, size_t count) {
    ws[slot].ensure(count * sizeof(T));
    return ws[slot].as<T>();
  }
  // Pinned host block of at least `bytes` for small device->host reads.
  void* host_pinned(size_t bytes);
  // Default-priority stream used only to capture HIP graphs.
  hipStream_t capture();
  // Replays `batch` (capturing `body(stream)` first when `key` changed) on
  // the graph's stream.
  template <class F>
  void replay(gx::CapturedBatch& batch, uint64_t key, F body) {
    if (!batch.exec || batch.key != key) {
      batch.reset();
      hipStream_t cs = capture();
      GX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      body(cs);
      hipGraph_t graph;
      GX_HIP(hipStreamEndCapture(cs, &graph));
      hipError_t e = hipGraphInstantiate(&batch.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      GX_HIP(e);
      batch.key = key;
    }
    GX_HIP(hipGraphLaunch(batch.exec, stream));
  }
};

namespace gx {

bool profiling_enabled();
// This device's address of the static scratch array of gx_cc_noalloc
// (resolved with hipGetSymbolAddress once per device, then cached).
unsigned char* noalloc_scratch();
// Value of a tuning knob (gx_tune_set) for a graph of n vertices and m
// edges: the value set, or the built-in default for that size when unset.
int64_t tune_effective(const char* key, int64_t n, int64_t m);
// Default of "pr.task_edges": m / 1024 within [2^19, 2^20].
int64_t pr_task_edges_default(int64_t n, int64_t m);
// Degree above which a frontier vertex is expanded by the whole grid.
// Frontier vertices of higher degree are expanded by the whole grid (default
// 128: a workgroup batch of 256 frontier vertices then holds at most 32K
// edges, so short power-law frontiers still spread over the grid).
inline uint32_t hub_degree() { return static_cast<uint32_t>(tune_effective("frontier.hub_degree", 0, 0)); }

// Accumulates the device time of the launches of one kernel with HIP events
// (enabled by gx_set_profiling).  Events are recorded on the graph's stream,
// the stream the kernel is launched on.
struct KernelTimer {
  hipStream_t s;
  bool on;
  std::vector<hipEvent_t> evs;
  explicit KernelTimer(hipStream_t st) : s(st), on(profiling_enabled()) {}
  ~KernelTimer() {
    for (auto e : evs) (void)hipEventDestroy(e);
  }
  void begin() {
    if (!on) return;
    hipEvent_t e;
    GX_HIP(hipEventCreate(&e));
    GX_HIP(hipEventRecord(e, s));
    evs.push_back(e);
  }
  void end() { begin(); }
  // Sum of (end - begin) over the first `limit` recorded pairs (all by
  // default); the stream must be synchronised.
  double total_ms(uint32_t limit = 0xffffffffu) {
    double t = 0;
    for (size_t i = 0; i + 1 < evs.size() && i / 2 < limit; i += 2) {
      float ms = 0;
      GX_HIP(hipEventElapsedTime(&ms, evs[i], evs[i + 1]));
      t += ms;
    }
    return t;
  }
  uint32_t count() const { return static_cast<uint32_t>(evs.size() / 2); }
};

// Wall-clock of a whole algorithm region on the graph's stream.
struct RegionTimer {
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  explicit RegionTimer(hipStream_t st) : s(st) {
    GX_HIP(hipEventCreate(&a));
    GX_HIP(hipEventCreate(&b));
    GX_HIP(hipEventRecord(a, s));
  }
  ~RegionTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  double stop() {
    GX_HIP(hipEventRecord(b, s));
    GX_HIP(hipEventSynchronize(b));
    float ms = 0;
    GX_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
};

inline void check_graph(const gx_graph* g) {
  GX_REQUIRE(g != nullptr, "graph handle is NULL");
}
inline void check_vertex(const gx_graph* g, uint32_t v, const char* what) {
  if (static_cast<int64_t>(v) >= g->n)
    fail(GX_ERR_INVALID, std::string(what) + " " + std::to_string(v) + " out of range (n = " +
                             std::to_string(g->n) + ")");
}
inline void check_ptr(const void* p, const char* what) {
  if (!p) fail(GX_ERR_INVALID, std::string(what) + " is NULL");
}

inline unsigned grid_for(int64_t items, int block, int64_t max_blocks = 1 << 20) {
  int64_t b = (items + block - 1) / block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return static_cast<unsigned>(b);
}

// Device-side helpers exported by graph_dev.hip for other translation units.
void ensure_outdeg(gx_graph* g);    // g->outdeg_buf = out-degree per vertex
void build_transpose(gx_graph* g);  // CSC + permutation (deterministic)
void build_degree_order(gx_graph* g);  // g->degree_order (relabelled transpose)
void build_relabeled(gx_graph* g);     // g->relab (degree-ordered CSR, uint32 weights carried)
// out[v] = in[new_of_old[v]] for v < n (results back to the caller's ids).
void unpermute_u32(const uint32_t* in, const uint32_t* new_of_old, uint32_t* out, int64_t n, hipStream_t s);
// Source vertex of every CSR edge into src[m].
void edge_sources(const int32_t* row, int64_t n, int64_t m, uint32_t* src, hipStream_t s);
// Exclusive prefix sum of `count` uint32 values into uint32 `out` (may alias).
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s);
// Largest out-degree of the graph (one reduction, then cached).
int64_t max_degree(gx_graph* g);
// Scans and radix sort of the build path (sort.hip): sum scan (exclusive),
// max scan (inclusive), both may run in place; stable ascending LSD radix
// sort of count < 2^32 keys over bits [0, end_bit) with optional uint32
// values carried along (inputs are not changed; in and out must differ).
void scan_u32_sum(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s);
void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s);
void radix_sort_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, int64_t count,
                    int end_bit, hipStream_t s);
// Ascending radix sort of count keys over bits [0, end_bit) (in is not changed).
void sort_keys_u64(const uint64_t* in, uint64_t* out, int64_t count, int end_bit, hipStream_t s);
// Out-edge chunks of the CSR: every vertex's adjacency cut into pieces of at
// most kEdgeChunk edges, (vertex, first edge) per chunk in vertex order;
// built once per graph and kept with it (bc.hip).
constexpr int kEdgeChunk = 256;
struct EdgeChunks {
  const uint32_t* vtx;
  const int32_t* begin;
  int64_t count;
};
EdgeChunks out_edge_chunks(gx_graph* g);
// Device->host copy of a few bytes through the graph's pinned block.
template <class T> T read_scalar(gx_graph* g, const T* dev) {
  T* h = static_cast<T*>(g->host_pinned(sizeof(T)));
  GX_HIP(hipMemcpyAsync(h, dev, sizeof(T), hipMemcpyDeviceToHost, g->stream));
  GX_HIP(hipStreamSynchronize(g->stream));
  return *h;
}
void read_bytes(gx_graph* g, void* host, const void* dev, size_t bytes);

}  // namespace gx
