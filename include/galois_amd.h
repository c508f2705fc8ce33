/*
 * galois_amd.h — C ABI of the MI355X (gfx950) lonestar analytics library.
 *
 * Plain C: integer status codes, plain pointers and sizes, no C++ or torch
 * types.  Every algorithm entry point takes a device graph (gx_graph) and
 * writes its result into a caller-owned DEVICE buffer; the library owns its
 * scratch (frontiers, buckets, per-source state) inside the gx_graph and
 * frees it with the graph.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the Galois source tree).  The reference's GPU path is the
 * LonestarGPU "skelapp" link-time ABI (libgpu/src/skelapp/skel.cu:12-32:
 * gg_main(CSRGraphTy& hg, CSRGraphTy& gg), output(), prog_opts) plus the
 * D-IrGL C-style personality kernels (e.g.
 * lonestar/analytics/distributed/k-core/kcore_push_cuda.h:1-80); the CPU path
 * is the lonestar app bodies under lonestar/analytics/cpu/.
 *
 * Conventions
 *  - Graph layout: CSR with int32 row offsets (n+1, row[0] = 0) and int32
 *    column indices (m); optional uint32 edge weights (m).  Vertex ids are
 *    0..n-1.  m must be < 2^31.
 *  - Distances/levels are uint32 with the reference infinity
 *    GX_DIST_INFINITY = UINT32_MAX/2 - 1 = 2147483646
 *    (lonestar/liblonestar/include/Lonestar/BFS_SSSP.h:31-32).
 *  - Calls on one gx_graph are not thread-safe; calls on different graphs are.
 *  - All work is enqueued on the graph's stream (gx_graph_set_stream; default
 *    the null stream) and every call returns after that stream has finished
 *    the work, so outputs are ready when the call returns.
 *  - Errors never exit(): a non-zero gx_status is returned and
 *    gx_last_error() holds a message (thread-local).
 */
#ifndef GALOIS_AMD_H
#define GALOIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GX_DIST_INFINITY 2147483646u /* BFS_SSSP.h:31-32 */
#define GX_BC_INFINITY 1073741823u   /* BetweennessCentrality.cpp:32 */

typedef enum gx_status {
  GX_OK = 0,
  GX_ERR_INVALID = 1,  /* bad argument: null pointer, size, out-of-range id */
  GX_ERR_HIP = 2,      /* a HIP runtime call or kernel launch failed       */
  GX_ERR_ALLOC = 3,    /* device or host allocation failed                 */
  GX_ERR_IO = 4,       /* file open/read/write or .gr format error         */
  GX_ERR_CAPACITY = 5, /* an internal queue bound would be exceeded        */
  GX_ERR_NO_DEVICE = 6 /* no usable gfx950 device                          */
} gx_status;

/* Graph flags. */
#define GX_GRAPH_SYMMETRIC 1u /* every edge (u,v) has its reverse (v,u)      */
#define GX_GRAPH_WEIGHTED 2u  /* uint32 weights present                      */
#define GX_GRAPH_FWEIGHTED 4u /* float32 weights present (adjoint PageRank)   */

typedef struct gx_graph gx_graph;

/* Per-call statistics (every field optional to read; pass NULL to skip). */
typedef struct gx_stats {
  uint32_t iterations;      /* levels / bucket phases / rounds / PR iterations */
  uint32_t launches;        /* kernel launches issued by the call              */
  uint64_t edges_visited;   /* edges scanned (the GTEPS numerator)             */
  uint64_t items;           /* vertices processed from worklists               */
  double residual;          /* PageRank: final L1 residual, computed in FP64   */
  double elapsed_ms;        /* device time of the algorithm region (hipEvent)  */
  double main_kernel_ms;    /* summed device time of the dominant kernel, when
                               gx_set_profiling(1) is on (else 0)             */
  uint32_t main_kernel_launches; /* launches of the dominant kernel          */
  uint32_t reserved;
  uint64_t main_kernel_bytes; /* algorithmic HBM bytes per launch of the
                                 dominant kernel (PageRank pull: per sweep),
                                 0 where an entry point does not state it     */
} gx_stats;

/* Host-side CSR as read from / written to a .gr file. */
typedef struct gx_host_csr {
  int64_t n, m;
  int32_t* row;        /* n+1 */
  int32_t* col;        /* m */
  void* edge_data;     /* m * edge_size bytes, or NULL */
  uint32_t edge_size;  /* bytes per edge datum (0 = none, 4 = uint32/float) */
  uint32_t version;    /* .gr version read (1 or 2); written files use 1 */
} gx_host_csr;

/* ---------------------------------------------------------------- errors */

/* Copy the calling thread's last error message into buf (always
 * NUL-terminated).  Returns the full message length. */
size_t gx_last_error(char* buf, size_t len);

/* Library/device query: 0 if a gfx950 device is usable. */
gx_status gx_init(int device);
/* When on, entry points time every launch of their dominant kernel with HIP
 * events and report it in gx_stats.main_kernel_ms.  Off by default. */
void gx_set_profiling(int on);

/* Tuning knobs (process-wide; no counterpart in the reference, whose
 * equivalents are compile-time constants such as CHUNK_SIZE/EDGE_TILE_SIZE,
 * bfs.cpp:84-86, ConnectedComponents.cpp:125).  Set by the autotuner
 * (galois_amd/autotune.py) or by hand; -1 restores the built-in default.
 *   "frontier.hub_degree"  frontier vertices with more edges than this are
 *                          expanded by the whole grid (BFS, SSSP, CC, k-core,
 *                          BC, push PageRank); default 128, range
 *                          [64, 2^30]
 *   "pr.hot_sources"       pull PageRank: in-edges from the first this-many
 *                          sources of the degree order are summed in a pass
 *                          of their own (their ranks stay L2-resident); 0 =
 *                          one pass (the default)
 *   "pr.blocking"          pull PageRank: 1 (default) = propagation-
 *                          blocked sweep for FP32 (streamed per-group
 *                          contributions, exact fixed-point row sums), 2 =
 *                          also for FP64 (FP64 atomic row sums, order
 *                          dependent in the last bit), 0 = merge-path pull
 *                          sweep for both (Kahan row sums in FP32)
 *   "pr.task_edges"        propagation-blocked sweep: a destination bin
 *                          with more in-edges than this is summed by
 *                          several workgroups (slices) that meet in a
 *                          global accumulator; read when a graph's layout
 *                          is built; default m/1024 within [2^19, 2^20],
 *                          range [2^12, 2^30]
 *   "bfs.visited_bitmap"   gx_bfs: 1 = test a visited bitmap before the
 *                          level word of every edge's target, 0 = not; the
 *                          default uses it from 2^25 vertices on
 *   "sssp.max_merge"       gx_sssp_delta: most consecutive buckets one
 *                          phase may cover (a phase merges buckets while the
 *                          previous one admitted fewer than 2^14 items);
 *                          1 = one bucket per phase as in the reference's
 *                          OBIM order; default 256, range [1, 2^16]
 *   "sssp.defer_hubs"      gx_sssp_delta: 1 = hubs admitted in a relax step
 *                          wait in a pending list expanded once the near
 *                          frontier has drained; 0 (default) = expanded in
 *                          the same step
 *   "sssp.hub_refine"      gx_sssp_delta: r > 0 defers hubs as above and
 *                          expands, per hub step, only the pending hubs
 *                          within 2^(r - 1) of the smallest pending hub
 *                          distance, so that hubs go out nearly in distance
 *                          order and most are expanded once; 0 = off;
 *                          hub selection joins a relax step whenever no near
 *                          item is closer than the hubs it takes; default 1
 *                          (one distance at a time) from 2^24 vertices on,
 *                          else 0; range [0, 32]
 *   "sssp.batch_width"     gx_sssp_batched: buckets per phase, shared by
 *                          the 64 sources of a wavefront; 0 = adaptive as
 *                          gx_sssp_delta's phases; default 64
 *   "frontier.relabel"     gx_sssp_delta, gx_bfs, gx_bfs_do (symmetric graphs)
 *                          and gx_kcore on graphs with hubs:
 *                          1 = run on a degree-ordered copy of the CSR
 *                          built once per graph (hot vertices' words
 *                          contiguous), 0 = on the CSR as given; default 1
 *                          from 2^24 vertices on
 *   "bc.chunk"             gx_bc_batched: edges per work item, the degree
 *                          bin boundary (an adjacency up to one chunk is one
 *                          wavefront's item, a longer one is split over
 *                          wavefronts that combine with atomics); default
 *                          256, range [64, 4096]
 *   "bc.grid"              gx_bc_batched: workgroups of each level sweep;
 *                          default 8192, range [256, 65536]
 * Unknown keys and out-of-range values are GX_ERR_INVALID.
 * gx_tune_get returns the value set (-1 when unset); gx_tune_effective the
 * value an entry point uses on a graph of n vertices and m edges: the value
 * set, else the built-in default for that size.  A slice may exceed
 * "pr.task_edges" by up to 511 edges: it is closed at the first job that
 * reaches the limit, and a job holds at most 512 edges. */
gx_status gx_tune_set(const char* key, int64_t value);
gx_status gx_tune_get(const char* key, int64_t* value);
gx_status gx_tune_effective(const char* key, int64_t n, int64_t m, int64_t* value);

/* ---------------------------------------------------------- graph lifecycle
 * Replaces CSRGraph::copy_to_gpu (libgpu/src/csr_graph.cu:274-293) and the
 * skelapp ownership of hg/gg (libgpu/src/skelapp/skel.cu:45-75). */

/* Copy a host CSR to the device (library-owned copy). */
gx_status gx_graph_create(int64_t n, int64_t m, const int32_t* row,
                          const int32_t* col, const void* weights,
                          uint32_t flags, gx_graph** out);
/* Wrap caller-owned device arrays (borrowed: not freed by gx_graph_free). */
gx_status gx_graph_create_device(int64_t n, int64_t m, int32_t* row_dev,
                                 int32_t* col_dev, void* weights_dev,
                                 uint32_t flags, gx_graph** out);
gx_status gx_graph_free(gx_graph* g);
gx_status gx_graph_info(const gx_graph* g, int64_t* n, int64_t* m,
                        uint32_t* flags);
/* hipStream_t passed as void* to keep HIP types out of the ABI. */
gx_status gx_graph_set_stream(gx_graph* g, void* stream);
/* Device pointers of the CSR / CSC arrays (any out-pointer may be NULL). */
gx_status gx_graph_device_arrays(const gx_graph* g, const int32_t** row,
                                 const int32_t** col, const void** weights,
                                 const int32_t** trow, const int32_t** tcol,
                                 const int32_t** tperm);
/* Copy the CSR back to host arrays sized n+1 / m / m*4 (any may be NULL). */
gx_status gx_graph_download(const gx_graph* g, int32_t* row, int32_t* col,
                            void* weights);
/* Copy the CSC (transpose) back: trow n+1, tcol m, tperm m (CSC slot -> CSR
 * edge index).  Builds it first if absent. */
gx_status gx_graph_download_transpose(gx_graph* g, int32_t* trow,
                                      int32_t* tcol, int32_t* tperm);

/* Build the transposed graph (CSC: in-edges) on the device.  Within each
 * column, sources are in ascending order and ties keep CSR edge order, so the
 * result is deterministic; equals graph-convert -transpose
 * (tools/graph-convert/graph-convert.cpp:1831-1880) and replaces
 * LC_CSR_CSC_Graph::constructIncomingEdges
 * (libgalois/include/galois/graphs/LC_CSR_CSC_Graph.h:140-244). */
gx_status gx_graph_build_transpose(gx_graph* g);

/* ------------------------------------------------------------- .gr files
 * Replaces FileGraph::fromMem (libgalois/src/FileGraph.cpp:44-53,104-141)
 * and CSRGraph::readFromGR (libgpu/src/csr_graph.cu:172-261). */
gx_status gx_gr_read(const char* path, gx_host_csr* out);
gx_status gx_gr_write(const char* path, const gx_host_csr* g);
void gx_host_csr_free(gx_host_csr* g);

/* ------------------------------------------------------------ generators
 * Seeded, counter-based (identical on every host and on the device).
 * RMAT: scale, edge_factor, (a,b,c) with d = 1-a-b-c; self-loops and
 * duplicate edges removed; symmetric = add the reverse of every edge.
 * weight_kind: 0 none, 1 uint32 uniform [wlo,whi] per undirected pair,
 * 2 float32 lognormal(0,1) per directed edge. */
gx_status gx_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                      double c, uint64_t seed, int symmetric, int weight_kind,
                      uint32_t wlo, uint32_t whi, gx_graph** out);
/* Road-like grid: side x side lattice, 4-neighbour edges, each undirected
 * edge dropped with probability drop_ppm/1e6, plus shortcut_ppm/1e6 * edges
 * random long-range undirected shortcuts; symmetric; uint32 weights
 * uniform [wlo,whi]. */
gx_status gx_gen_grid(uint32_t side, uint32_t drop_ppm, uint32_t shortcut_ppm,
                      uint64_t seed, uint32_t wlo, uint32_t whi,
                      gx_graph** out);
/* Disjoint union of graphs (vertex ids of graph i offset by the vertex count
 * of graphs 0..i-1); the result owns copies.  Used to run many uneven graphs
 * in one launch. */
gx_status gx_graph_union(gx_graph* const* parts, uint32_t count,
                         gx_graph** out, int64_t* vertex_offsets /* count+1 */);

/* Host copy of the source picker of lonestar/liblonestar/include/Lonestar/
 * Utils.h:12-31 (mt19937 seeded 27491095, uniform over [0,n-1], rejecting
 * degree-0 vertices).  row is the host CSR row array. */
gx_status gx_pick_sources(int64_t n, const int32_t* row, uint32_t count,
                          uint32_t seed, uint32_t* out);

/* Host-only degree-distribution test that the apps' AutoAlgo choice uses
 * (isApproximateDegreeDistributionPowerLaw, lonestar/liblonestar/include/
 * Lonestar/Utils.h:37-57, called by SSSP.cpp:410-419, bfsDirectionOpt.cpp:
 * 469-478 and BetweennessCentrality.cpp:110-122): *power_law = 1 when the
 * integer average degree m/n is at least 10 and, over the degrees of 1000
 * vertices drawn by gx_pick_sources (seed 27491095), mean / 1.25 > median.
 * row is the host CSR row array (n+1, row[n] == m). */
gx_status gx_degree_power_law(int64_t n, int64_t m, const int32_t* row,
                              int* power_law);

/* ------------------------------------------------------------ algorithms
 * Outputs are DEVICE pointers owned by the caller. */

/* Level-synchronous BFS (top-down, degree-binned expansion).
 * level[v] = hops from src, GX_DIST_INFINITY if unreached.
 * Replaces bfs gg_main (lonestar/analytics/gpu/bfs/bfs.cu:349-362) and
 * bfs-cpu syncAlgo (lonestar/analytics/cpu/bfs/bfs.cpp:245-293). */
gx_status gx_bfs(gx_graph* g, uint32_t src, uint32_t* level, gx_stats* st);

/* Direction-optimising BFS (push while the frontier is small, pull over
 * in-edges with a frontier bitmap otherwise; alpha/beta as in
 * lonestar/analytics/cpu/bfs/bfsDirectionOpt.cpp:68-77, 0 = defaults 15/18).
 * Writes levels (same values as gx_bfs) and, when parent != NULL, a valid
 * BFS parent per reached vertex (parent[src] = src, unreached = UINT32_MAX).
 * Replaces syncDOAlgo (bfsDirectionOpt.cpp:205-342). */
gx_status gx_bfs_do(gx_graph* g, uint32_t src, uint32_t alpha, uint32_t beta,
                    uint32_t* level, uint32_t* parent, gx_stats* st);

/* Delta-stepping SSSP on uint32 weights; bucket = dist >> shift (the
 * reference's -delta option is this shift).  dist[v] = shortest distance,
 * GX_DIST_INFINITY if unreached.  Replaces sssp gg_main
 * (lonestar/analytics/gpu/sssp/sssp.cu:407-432) and deltaStepAlgo
 * (lonestar/analytics/cpu/sssp/SSSP.cpp:119-173). */
gx_status gx_sssp_delta(gx_graph* g, uint32_t src, uint32_t shift,
                        uint32_t* dist, gx_stats* st);
/* Shortest paths from many sources at once: dist (device uint32[count * n],
 * source-major: dist[i * n + v] for sources[i]) equals gx_sssp_delta from
 * each source (GX_DIST_INFINITY where unreached).  Sources run 64 per
 * wavefront (one lane each) with a shared delta-bucket order (bucket =
 * dist >> shift, phases may merge buckets as gx_sssp_delta's do); the round
 * loop is device-driven and captured in HIP graphs.  The reference runs one
 * source per execution of SSSP.cpp:119-173; no counterpart for the batch.
 * sources: HOST array. */
gx_status gx_sssp_batched(gx_graph* g, const uint32_t* sources, uint32_t count, uint32_t shift, uint32_t* dist,
                          gx_stats* st);

/* Connected components of a symmetric graph; label[v] = minimum vertex id
 * of v's component (the canonical result of the reference union-find,
 * libgalois/include/galois/UnionFind.h:108-122, and LabelProp,
 * ConnectedComponents.cpp:183-230).  Replaces cc gg_main
 * (lonestar/analytics/gpu/connected-components/cc.cu:561-655). */
gx_status gx_cc(gx_graph* g, uint32_t* label, gx_stats* st);
/* gx_cc without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device or pinned host memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * labels as gx_cc; the kernels' scratch is a static device array of
 * 64 MiB + 1.5 KiB that each GPU holds until the process ends, so the graph
 * may have at most 2^22 vertices (GX_ERR_CAPACITY above that).  Calls that
 * overlap on one GPU share that array and race: run them on one stream or
 * serialise them.  No timing, no statistics. */
gx_status gx_cc_noalloc(gx_graph* g, uint32_t* label);

/* k-core of a symmetric graph by synchronous cascade peeling.
 * degree[v] ends as the number of alive neighbours (alive: degree >= k),
 * exactly the reference's currentDegree; alive (nullable) gets 0/1; the
 * alive count is returned in *alive_count (nullable).  Replaces
 * syncCascadeKCore (lonestar/analytics/cpu/k-core/kcore.cpp:124-159) and the
 * D-IrGL KCoreStep1/2 kernels (kcore_push_cuda.py:10-87). */
gx_status gx_kcore(gx_graph* g, uint32_t k, uint32_t* degree, uint8_t* alive,
                   uint64_t* alive_count, gx_stats* st);

/* PageRank, pull over the transpose (built on demand).
 * Semantics (normalised): x0 = 1/n; x' = (1-alpha)/n + alpha*(sum_{u->v}
 * x[u]/outdeg(u) + D/n), D = sum of x over dangling (outdeg 0) vertices;
 * stop when ||x'-x||_1 < tol (FP64 reduction) or after max_iter iterations.
 * precision: 0 = FP32 ranks, 1 = FP64 ranks.  FP32 runs by default
 * through the propagation-blocked sweep, whose row sums are exact (every
 * contribution rounded to 64-bit fixed point with resolution 2^-62 and
 * summed with integer atomics, so the sum is order independent and more
 * accurate than a compensated FP32 sum), and x' = base + alpha*sum is
 * evaluated in FP32 with error-free transformations; with the knob
 * "pr.blocking" = 0 it runs the merge-path pull with Kahan-compensated FP32
 * row sums.  In both, the residual and the dangling mass D are reduced in
 * FP64 (more accurate than an FP32 reduction and read once per iteration).
 * pr: device float[n] (precision 0) or double[n] (precision 1).
 * Replaces computePRTopological (PageRank-pull.cpp:196-251) with the
 * normalised, dangling-aware semantics. */
#define GX_PR_FP32 0
#define GX_PR_FP64 1
gx_status gx_pagerank_pull(gx_graph* g, double alpha, double tol,
                           uint32_t max_iter, int precision, void* pr,
                           gx_stats* st);
/* Shape of the propagation-blocked layout a graph's FP32 pull PageRank runs
 * on (built by the first such call and kept with the graph; all zero
 * before): destination bins (= source chunks) of 16384 vertices, (source,
 * bin) groups, phase-1 pieces, phase-2 tasks, bins summed in slices, phase-2
 * jobs per length class, and the slice size in edges it was built with
 * ("pr.task_edges").  No reference counterpart. */
typedef struct gx_pb_layout_info {
  int64_t bins, groups, pieces, tasks, sliced_bins;
  int64_t jobs_long, jobs_mid, jobs_short;
  int64_t task_edges;
  int32_t built, reserved;
} gx_pb_layout_info;
gx_status gx_pagerank_layout_info(const gx_graph* g, gx_pb_layout_info* info);

/* Reference-semantics pull PageRank (unnormalised, no dangling
 * redistribution), FP32, the two reference variants:
 * variant 0 = residual (PageRank-pull.cpp:132-190; value starts 0, residual
 * starts 1-alpha, stop when no residual exceeds tol), variant 1 = Jacobi
 * topological (value starts 1/n; x' = alpha*sum x[u]/outdeg(u) +
 * (1-alpha)/n; stop when sum|x'-x| <= tol).  pr: device float[n]. */
gx_status gx_pagerank_pull_ref(gx_graph* g, int variant, float alpha,
                               float tol, uint32_t max_iter, float* pr,
                               gx_stats* st);

/* Residual-push PageRank, FP32 atomics, the synchronous residual push of
 * lonestar/analytics/cpu/pagerank/PageRank-push.cpp:103-188 (the GPU app
 * lonestar/analytics/gpu/pagerank/pagerank.cu:103-236 is its asynchronous
 * form): value starts 0 and residual 1-alpha; every round each active vertex
 * whose residual exceeds tol adds it to its value and pushes
 * alpha*residual/outdeg to its out-neighbours; a neighbour becomes active when
 * its residual crosses tol.  Converges to the unnormalised PageRank
 * x = (1-alpha) + alpha*sum x[u]/outdeg(u).  pr: device float[n]. */
gx_status gx_pagerank_push(gx_graph* g, float alpha, float tol,
                           uint32_t max_iter, float* pr, gx_stats* st);

/* Differentiable PageRank on a weighted directed graph.
 * P[u,v] = w_uv / sum_v' w_uv'; rows with no out-edge (or zero weight sum)
 * are dangling and spread uniformly; x = alpha*P^T x + alpha*(D/n) +
 * (1-alpha)*t, D = dangling mass, solved by Jacobi iteration to
 * sum|x'-x| < tol or max_iter (x0 = 1/n).  weights: device float[m] in CSR
 * edge order (NULL = the graph's own float weights); teleport: device
 * float[n].  precision GX_PR_FP32 (compensated sums) or GX_PR_FP64; x is a
 * device float[n] or double[n] accordingly.
 * Extends computePRTopological (PageRank-pull.cpp:196-251) to edge weights
 * and a teleport vector; no reference counterpart for the gradient. */
gx_status gx_pagerank_weighted(gx_graph* g, const float* weights, const float* teleport, double alpha,
                               double tol, uint32_t max_iter, int precision, void* x, gx_stats* st);
/* Adjoint of gx_pagerank_weighted for a loss L with dL/dx = c (device
 * float[n]): solves lambda = c + alpha*(P lambda + (sum(lambda)/n) on dangling
 * rows) by the transposed power iteration on the same CSR, then writes
 * grad_w[e=(u,v)] = alpha*x[u]*(lambda[v] - (P lambda)[u]) / s_u (CSR edge
 * order) and grad_t = (1-alpha)*lambda.  x is the converged forward vector
 * (the only forward state the backward pass needs); x, grad_w, grad_t are
 * float or double per `precision`. */
gx_status gx_pagerank_adjoint(gx_graph* g, const float* weights, const void* x, const float* c, double alpha,
                              double tol, uint32_t max_iter, int precision, void* grad_w, void* grad_t,
                              gx_stats* st);

/* Brandes betweenness centrality, one source per pass (plain).
 * bc[v] (device double[n]) = sum over the given sources s != v of the
 * dependency delta_s(v); directed convention, not halved (as
 * LevelStructs.h:142-185 / OuterStructs.h:41-102).  Path counts sigma in
 * FP32, dependencies in FP64.  sources: HOST array. */
gx_status gx_bc(gx_graph* g, const uint32_t* sources, uint32_t count,
                double* bc, gx_stats* st);
/* Same result, all sources batched: 64 sources per wavefront (one lane per
 * source), level loops captured in a HIP graph.  On a symmetric graph each
 * group of 64 sources sweeps only the vertex span of its sources' connected
 * components (labels computed on the first call and kept with the graph), so
 * a union of many graphs (gx_graph_union) costs each group its own graphs.
 * sources: HOST array. */
gx_status gx_bc_batched(gx_graph* g, const uint32_t* sources, uint32_t count,
                        double* bc, gx_stats* st);

/* ------------------------------------------------- D-IrGL personality
 * The per-operator GPU kernels that D-IrGL host code launches on one host's
 * partition (lonestar/analytics/distributed/bfs/bfs_push_cuda.h:103-140,
 * k-core/kcore_push_cuda.h:133-167).  Node fields live on the device in a
 * partition context; every operator runs over local ids [begin, end) and
 * returns its reductions through out-parameters (the reference's
 * DGAccumulator / DGReduceMax arguments).  The multi-host mirror exchange
 * (batch_get/set/min_*_cuda over master/mirror lists) is gluon's and not part
 * of this one-GPU build.  A partition is not thread-safe; each call returns
 * after its work has finished. */
typedef struct gx_part gx_part;

/* get_CUDA_context + init_CUDA_context (libgluon/include/galois/cuda/
 * HostDecls.h:98-99). */
gx_status gx_part_create(int device, gx_part** out);
/* reset_CUDA_context / context teardown (HostDecls.h:102). */
gx_status gx_part_free(gx_part* p);
/* load_graph_CUDA (HostDecls.h:100-101) from the MarshalGraph fields
 * (HostDecls.h:38-54): host arrays row_start[n+1] and edge_dst[m] (uint32, the
 * reference's index_type), edge_data[m] (uint32 weights, needed by SSSP push;
 * NULL = none), node_gid[n] = global id of each local node (the MarshalGraph
 * node_data; NULL = identity), the master range [begin_master, begin_master +
 * num_owned) and numNodesWithEdges.  The arrays are validated (offsets
 * monotone from 0 to m, destinations < n) before upload.  Fields and bitsets
 * start at zero. */
gx_status gx_part_load(gx_part* p, uint64_t n, uint64_t m, const uint32_t* row_start,
                       const uint32_t* edge_dst, const uint32_t* edge_data, const uint32_t* node_gid,
                       uint32_t num_owned, uint32_t begin_master, uint32_t num_nodes_with_edges);
/* Ranges of the reference's _allNodes_cuda / _masterNodes_cuda /
 * _nodesWithEdges_cuda variants. */
#define GX_RANGE_ALL 0
#define GX_RANGE_MASTERS 1
#define GX_RANGE_WITH_EDGES 2
gx_status gx_part_range(const gx_part* p, int which, uint32_t* begin, uint32_t* end);

/* Node fields (uint32 each; flag is uint8 in the reference, 0/1 here). */
#define GX_FIELD_DIST_CURRENT 0   /* bfs_push NodeData (bfs_push.cpp:74-77) */
#define GX_FIELD_DIST_OLD 1
#define GX_FIELD_CURRENT_DEGREE 2 /* kcore_push NodeData (kcore_push.cpp:69) */
#define GX_FIELD_TRIM 3
#define GX_FIELD_FLAG 4
#define GX_FIELD_COMP_CURRENT 5   /* cc_push NodeData (cc_push.cpp:63-66) */
#define GX_FIELD_COMP_OLD 6
#define GX_FIELD_PR_VALUE 7       /* pagerank_push NodeData (pagerank_push.cpp:69-74); */
#define GX_FIELD_PR_NOUT 8        /* value, residual and delta are float bit patterns */
#define GX_FIELD_PR_RESIDUAL 9
#define GX_FIELD_PR_DELTA 10
#define GX_FIELD_BC_LENGTH 11     /* bc_level NodeData (bc_level.cpp:86-92): current_length, */
#define GX_FIELD_BC_DEPENDENCY 12 /* dependency and betweeness_centrality (float bit */
#define GX_FIELD_BC_VALUE 13      /* patterns), num_shortest_paths (FP64: read with */
#define GX_FIELD_BC_PATHS 14      /* gx_part_read_paths; bitset only otherwise) */
/* get/set/add/min_node_<field>_cuda (bfs_push_cuda.h:11-17); *changed = 1 when
 * min lowered the value. */
gx_status gx_part_get(const gx_part* p, int field, uint32_t lid, uint32_t* value);
gx_status gx_part_set(gx_part* p, int field, uint32_t lid, uint32_t value);
gx_status gx_part_add(gx_part* p, int field, uint32_t lid, uint32_t value);
gx_status gx_part_min(gx_part* p, int field, uint32_t lid, uint32_t value, int* changed);
/* batch_reset_node_<field>_cuda(ctx, begin, end, v) (bfs_push_cuda.h:55-56). */
gx_status gx_part_fill(gx_part* p, int field, uint32_t begin, uint32_t end, uint32_t value);
/* Field values of [begin, end) to a host array (the per-node get loop of
 * the reference's output stage, bfs_push.cpp:375, in one copy). */
gx_status gx_part_read(const gx_part* p, int field, uint32_t begin, uint32_t end, uint32_t* host);
/* get_bitset_<field>_cuda (ceil(n/64) words to host) and
 * bitset_<field>_reset_cuda, whole ([0, n)) or [begin, end)
 * (bfs_push_cuda.h:6-10).  Operators set bits on the field they change:
 * BFS/SSSP on dist_current, CC on comp_current, PageRank on nout (init)
 * and residual, InitializeGraph2 on current_degree, KCoreStep1 on trim. */
gx_status gx_part_bitset_get(const gx_part* p, int field, uint64_t* words);
gx_status gx_part_bitset_reset(gx_part* p, int field, uint32_t begin, uint32_t end);

/* BFS push operators (bfs_push.cpp:95-330).
 * InitializeGraph_cuda: dist_current = dist_old = (gid == src_gid ? 0 : infinity).
 * FirstItr_BFS_cuda: every node of the range copies dist_current to dist_old
 *   and pushes dist_current + 1 to its out-neighbours (atomic min, bitset).
 * BFS_cuda: nodes with dist_old > dist_current count as active; those with
 *   dist_current < priority also push; *work_items counts pushed edges.
 * BFSSanityCheck_cuda: *sum = nodes with dist_current < infinity, *max_dist =
 *   their largest distance. */
gx_status gx_part_bfs_init(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t src_gid);
gx_status gx_part_bfs_first(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_bfs(gx_part* p, uint32_t begin, uint32_t end, uint32_t priority,
                      uint32_t* active_vertices, uint32_t* work_items);
gx_status gx_part_bfs_check(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t* sum,
                            uint32_t* max_dist);
/* SSSP push operators (sssp/sssp_push.cpp:95-360, sssp_push_cuda.h:103-128):
 * as BFS push with dist_current + edge weight pushed; the check also returns
 * the sum of the reached distances (dg_avg). */
gx_status gx_part_sssp_init(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t src_gid);
gx_status gx_part_sssp_first(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_sssp(gx_part* p, uint32_t begin, uint32_t end, uint32_t priority,
                       uint32_t* active_vertices, uint32_t* work_edges);
gx_status gx_part_sssp_check(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t* sum,
                             uint64_t* dist_sum, uint32_t* max_dist);
/* CC push operators (connected-components/cc_push.cpp:85-300,
 * cc_push_cuda.h:103-130).
 * InitializeGraph_cuda: comp_current = comp_old = global id.
 * FirstItr_ConnectedComp_cuda: every node of the range pushes its label.
 * ConnectedComp_cuda: nodes with comp_old > comp_current push their label
 *   (atomic min, bitset comp_current); one active per pushed edge.
 * ConnectedCompSanityCheck_cuda: *components = nodes whose label is their
 *   own global id. */
gx_status gx_part_cc_init(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_cc_first(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_cc(gx_part* p, uint32_t begin, uint32_t end, uint32_t* active_vertices);
gx_status gx_part_cc_check(gx_part* p, uint32_t begin, uint32_t end, uint64_t* components);
/* PageRank push operators (pagerank/pagerank_push.cpp:110-430), FP32 fields;
 * alpha is the reference's reset probability (1 - 0.85, pagerank_push.cpp:68).
 * ResetGraph_cuda: value = nout = residual = delta = 0.
 * InitializeGraph_cuda: residual = alpha, nout += out-degree (bitset nout).
 * PageRank_delta_cuda: a positive residual moves into value; above the
 *   tolerance it also sets delta = residual * (1 - alpha) / nout.
 * PageRank_cuda: a node with positive delta (counted active) clears it and
 *   adds it to every out-neighbour's residual (FP32 atomics: the sum order,
 *   and so the last bits, vary run to run as in the reference; bitset
 *   residual).
 * PageRankSanity_cuda: extremes and sums of value and residual, residuals
 *   above the tolerance (sums in FP64). */
typedef struct gx_pr_sanity {
  float max_value, min_value, max_residual, min_residual;
  double sum_value, sum_residual;
  uint64_t residual_over_tolerance;
} gx_pr_sanity;
gx_status gx_part_pr_reset(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_pr_init(gx_part* p, uint32_t begin, uint32_t end, float alpha);
gx_status gx_part_pr_delta(gx_part* p, uint32_t begin, uint32_t end, float alpha, float tolerance);
gx_status gx_part_pr(gx_part* p, uint32_t begin, uint32_t end, uint32_t* active_vertices);
gx_status gx_part_pr_check(gx_part* p, uint32_t begin, uint32_t end, float tolerance, gx_pr_sanity* out);
/* BC level operators (betweennesscentrality/bc_level.cpp:150-560,
 * bc_level_cuda.h:223-275).
 * InitializeGraph_cuda: centrality, path counts, dependencies zero.
 * InitializeIteration_cuda: the source (global id) at length 0 with one
 *   path, every other node at infinity with none; dependency zero.
 * ForwardPass_cuda(round): nodes at length round push round + 1; a neighbour
 *   reached at round + 1 adds the node's path count (FP64); *dga counts those
 *   updates (bitsets current_length, num_shortest_paths).
 * MiddleSync_cuda: reached nodes marked in the num_shortest_paths bitset.
 * BackwardPass_cuda(round): nodes at length round take dependency =
 *   paths * sum over successors at round + 1 of (1 + dependency) / paths
 *   (FP32 field, per-node sum in a fixed order; bitset dependency).
 * BC_cuda: a positive dependency adds to betweeness_centrality.
 * Sanity_cuda: sum, maximum and minimum centrality. */
gx_status gx_part_bc_init(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_bc_iter(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity, uint64_t src_gid);
gx_status gx_part_bc_forward(gx_part* p, uint32_t begin, uint32_t end, uint32_t round, uint32_t* dga);
gx_status gx_part_bc_middle(gx_part* p, uint32_t begin, uint32_t end, uint32_t infinity);
gx_status gx_part_bc_backward(gx_part* p, uint32_t begin, uint32_t end, uint32_t round);
gx_status gx_part_bc_accumulate(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_bc_check(gx_part* p, uint32_t begin, uint32_t end, float* sum, float* max_bc, float* min_bc);
gx_status gx_part_read_paths(const gx_part* p, uint32_t begin, uint32_t end, double* host);
/* k-core push operators (kcore_push.cpp:93-360: InitializeGraph2, InitializeGraph1,
 * KCoreStep2, KCoreStep1, KCoreSanityCheck).
 * InitializeGraph1_cuda: flag = 1, trim = current_degree = 0.
 * InitializeGraph2_cuda: each out-edge adds 1 to its destination's
 *   current_degree (bitset current_degree).
 * KCoreStep1_cuda: a node with flag and current_degree < k dies (flag = 0,
 *   counted active) and adds 1 to each out-neighbour's trim (bitset trim).
 * KCoreStep2_cuda: a live node subtracts its trim from current_degree; trim = 0.
 * KCoreSanityCheck_cuda: *alive = nodes with flag set. */
gx_status gx_part_kcore_init1(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_kcore_init2(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_kcore_step1(gx_part* p, uint32_t begin, uint32_t end, uint32_t k, uint32_t* active_vertices);
gx_status gx_part_kcore_step2(gx_part* p, uint32_t begin, uint32_t end);
gx_status gx_part_kcore_check(gx_part* p, uint32_t begin, uint32_t end, uint64_t* alive);

#ifdef __cplusplus
}
#endif
#endif /* GALOIS_AMD_H */
