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
//    each round relaxes all of them with the load-balanced expansion of
//    frontier.hpp and atomicMin on the distance word; a vertex enters a
//    round's queue once (stamp word per vertex).  A vertex is expanded only
//    if its distance changed since it was last expanded (`done` word), the
//    GPU form of the reference's stale-item test (SSSP.cpp:138).
//  - far pile: vertices improved to a later bucket; each vertex is in the far
//    pile at most once (farmark word), so the pile is bounded by n.  A
//    running minimum of the far buckets (every improvement into a later
//    bucket, and every item a split keeps) names the next bucket; when the
//    near frontier drains, one pass moves that bucket into the near frontier
//    and keeps the rest.  The running minimum can name a bucket whose items
//    have meanwhile moved nearer; that split then moves nothing and the next
//    one uses the exact minimum of the items it kept.
//  - Merged buckets: a step covers the buckets [bucket, bucket + width).
//    width starts at 1 and doubles at an advance when the phase just ended
//    admitted fewer than kMergeLow items (up to the knob sssp.max_merge),
//    and halves when it admitted more than kMergeHigh: on road-like graphs
//    a delta-64 bucket holds a thin shell of a few thousand vertices, so
//    merging cuts the number of dependent steps, at the cost of a few
//    vertices expanded before their distance is final (they are expanded
//    again when it improves, as in any label-correcting order).  Power-law
//    graphs keep width 1 (their buckets hold millions of items).
//  - Hubs (degree above frontier.hub_degree) admitted in a relax step are
//    expanded over the whole grid in the same step.  With the knob
//    sssp.defer_hubs they instead join a pending list (once per list, mark
//    word per vertex) that is expanded in a step of its own when the near
//    frontier has drained, so a hub whose distance improves several times
//    inside a bucket is expanded once per drain rather than once per
//    improvement.  With sssp.hub_refine a step expands only the pending
//    hubs within a distance slice above the smallest pending hub distance
//    (one distance by default) and keeps the rest pending, so hubs go out
//    nearly in distance order, almost all at their final distance, and a
//    hub's hundreds of thousands of edges are relaxed about once instead
//    of once per improvement (RMAT-26: 1.84 -> 1.10 relaxations per edge).
//    The hub selection joins a relax step whenever the smallest pending hub
//    distance is at or below the smallest distance in that step's near
//    queue (hubs above it wait; positive weights mean nothing left to
//    expand can lower the hubs taken), and otherwise runs in a step of its
//    own once the near frontier drains.  The extra steps pay on large
//    power-law graphs only, so the default enables it from 2^24 vertices
//    on.
//  - Locality: on graphs with hubs (knob frontier.relabel) the kernels run
//    on the degree-ordered copy of the CSR (gx_graph::relab): the hubs and
//    the vertices most often relaxed get the smallest ids, so their
//    distance words share a few megabytes that stay cached, instead of each
//    relaxation moving a line of a 268 MB array (RMAT-26).  The source is
//    mapped in and the distances mapped back at the end.
//  - The loop is driven on the device: every step is the same fixed kernel
//    sequence (step; work, which is either the relax, with the stale-item
//    filter fused in as the expansion's admission test, or the far split, as
//    the step's mode says; hub layout; hubs) whose kernels read the step's
//    mode from a device state word and return at once when it is not
//    theirs; the step kernel also closes the previous step.  On graphs
//    without hubs the two hub launches are left out.  Steps are captured
//    once into a HIP graph (kStepsPerGraph steps and a closing kernel) and
//    replayed until the state says done, so the host synchronises once per
//    graph replay instead of once per round.
#include <algorithm>

#include "frontier.hpp"
#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kInf = GX_DIST_INFINITY;
constexpr int kGrid = 256 * 8;
// Graphs without hubs (maximum degree at or below the hub threshold, such as
// road-like grids) have small frontiers per step and many steps; a quarter
// of the workgroups per launch costs less there (2048-grid SSSP 26.4 ->
// 21.7 ms, BFS 2.03 -> 1.57 ms), while RMAT frontiers need the full grid.
constexpr int kGridNoHubs = 512;
constexpr int kStepsPerGraph = 16;
constexpr int kHubSelectGrid = 256;
// Bucket merging thresholds (items admitted per phase).
constexpr unsigned long long kMergeLow = 1ull << 14;
constexpr unsigned long long kMergeHigh = 1ull << 17;

enum Mode : unsigned { kIdle = 0, kRelax = 1, kAdvance = 2, kHubs = 3 };

struct SsspState {
  unsigned near_count[2];  // near queues (double buffered by q)
  unsigned far_count[2];   // far piles (double buffered by far_sel)
  unsigned hub_count;
  unsigned overflow;
  unsigned far_min;        // running minimum of the far buckets
  unsigned mode;
  unsigned done;
  unsigned q;              // near queue read this step
  unsigned far_sel;        // far pile appended to
  unsigned bucket;
  unsigned width;          // buckets merged into the current phase
  unsigned max_width;      // knob sssp.max_merge
  unsigned hub_gen;        // generation of the pending hub list
  unsigned hub_expand;     // hubs expanded this step (selected in a hub step, else 0)
  unsigned hq;             // pending hub list appended to (double buffered)
  unsigned hub_kept;       // pending hubs a hub step leaves for a later one
  uint32_t hub_min;        // smallest distance a pending hub had when it was listed
  uint32_t hub_min_next;   // the same over the hubs a hub step keeps
  uint32_t hub_limit;      // a hub step expands the pending hubs at or below this distance
  unsigned hub_step;       // this step selects and expands pending hubs
  uint32_t near_min;       // smallest distance in the near queue the step reads
  uint32_t near_min_next;  // the same for the near queue the step fills
  unsigned long long phase_items0;  // items admitted before the current phase
  unsigned round;
  unsigned phase;
  unsigned steps;
  unsigned long long edges;
  unsigned long long items;
  // Output of the current call, read by the captured step kernels.
  uint32_t* dist_out;
};

struct Bufs {
  const int32_t* row;
  const int32_t* col;
  const uint32_t* w;
  uint32_t* dist;
  uint32_t* stamp;
  uint32_t* farmark;
  uint32_t* done;
  uint32_t* hubmark;  // generation of the pending hub list a hub is in
  uint32_t* nearq[2];
  uint32_t* farq[2];
  uint32_t* hubq[2];  // pending hub lists (hq); without deferral hubq[0] holds the step's hubs
  uint32_t* hubx;     // hubs a hub step expands
  SsspState* S;
  uint32_t shift;
  unsigned cap;
  uint32_t* hub_off;
  uint32_t hub;  // frontier.hub_degree
  uint32_t defer;  // sssp.defer_hubs: pending hub list expanded in hub steps
  // sssp.hub_refine r: a hub step expands the pending hubs within
  // hub_span = 2^(r - 1) of the smallest pending hub distance; 0 = all.
  uint32_t hub_span;
};

__global__ void sssp_init_kernel(uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp,
                                 uint32_t* __restrict__ farmark, uint32_t* __restrict__ done,
                                 uint32_t* __restrict__ hubmark, int64_t n, uint32_t src) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) {
    dist[v] = (v == src) ? 0u : kInf;
    stamp[v] = 0xffffffffu;
    farmark[v] = 0xffffffffu;
    done[v] = kInf;
    hubmark[v] = 0xffffffffu;
  }
}

__global__ void sssp_reset_kernel(SsspState* S, uint32_t* nearq0, uint32_t src, uint32_t* dist,
                                  unsigned max_width) {
  if (threadIdx.x == 0) {
    S->dist_out = dist;
    S->near_count[0] = 1;
    S->near_count[1] = 0;
    S->far_count[0] = S->far_count[1] = 0;
    S->hub_count = 0;
    S->overflow = 0;
    S->far_min = 0xffffffffu;
    S->mode = kIdle;
    S->done = 0;
    S->q = 0;
    S->far_sel = 0;
    S->bucket = 0;
    S->width = 1;
    S->max_width = max_width;
    S->hub_gen = 0;
    S->hub_expand = 0;
    S->hq = 0;
    S->hub_kept = 0;
    S->hub_min = S->hub_min_next = S->hub_limit = 0xffffffffu;
    S->hub_step = 0;
    S->near_min = 0;  // the source
    S->near_min_next = 0xffffffffu;
    S->phase_items0 = 0;
    S->round = 0;
    S->phase = 0;
    S->steps = 0;
    S->edges = 0;
    S->items = 0;
    nearq0[0] = src;
  }
}

// Closes a step: the queue it filled becomes the input; after a split the
// far piles swap; after a step that expanded hubs the pending list holds
// the hubs it kept, under a new generation.
__device__ __forceinline__ void close_step(SsspState* S) {
  if (S->hub_step) {
    S->hq ^= 1;
    S->hub_count = S->hub_kept;  // kept hubs carry the new generation's mark
    S->hub_kept = 0;
    S->hub_expand = 0;
    S->hub_gen += 1;
    S->hub_min = S->hub_min_next;
    S->hub_step = 0;
  }
  if (S->mode == kRelax || S->mode == kHubs) {
    S->q ^= 1;
  } else if (S->mode == kAdvance) {
    S->far_count[S->far_sel] = 0;
    S->far_sel ^= 1;
    S->q ^= 1;
  }
  if (S->mode != kIdle) {
    S->near_min = S->near_min_next;
    S->near_min_next = 0xffffffffu;
  }
  S->mode = kIdle;
}

// Starts the hub part of a step: pending hubs at or below `limit` are
// selected (hub_select_kernel) and expanded (hubs_kernel).
__device__ __forceinline__ void begin_hub_step(SsspState* S, uint32_t limit) {
  S->hub_step = 1;
  S->hub_expand = 0;  // counted by hub_select_kernel
  S->hub_kept = 0;
  S->hub_min_next = 0xffffffffu;
  S->hub_limit = limit;
}

// Closes the previous step, then chooses this step's mode and clears the
// counters it appends to: relax while the near frontier has items, then
// expand the pending hubs, then advance.  An advance takes the running far
// minimum as the new first bucket and sets the phase's width from the items
// the last phase admitted; with far items left but no live far bucket
// (minimum unset) every far item is settled and the run is done.
__global__ void step_kernel(SsspState* S, uint32_t defer, uint32_t hub_span) {
  if (threadIdx.x != 0) return;
  close_step(S);
  if (S->done || S->overflow) return;
  S->round += 1;
  S->near_count[S->q ^ 1] = 0;
  if (!defer) S->hub_count = 0;  // hubs of a relax step are expanded in that step
  const uint32_t span_limit =
      (hub_span == 0 || S->hub_min >= 0xffffffffu - hub_span) ? 0xffffffffu : S->hub_min + (hub_span - 1u);
  if (S->near_count[S->q] > 0) {
    S->mode = kRelax;
    // With hub slices, the pending hubs at or below every near distance are
    // expanded in the same step as the near frontier: nothing left to
    // expand can still lower them (weights are positive).
    if (defer && hub_span != 0 && S->hub_count > 0 && S->hub_min <= S->near_min)
      begin_hub_step(S, min(span_limit, S->near_min));
  } else if (defer && S->hub_count > 0) {
    S->mode = kHubs;
    begin_hub_step(S, span_limit);
  } else if (S->far_count[S->far_sel] > 0 && S->far_min != 0xffffffffu) {
    S->mode = kAdvance;
    const unsigned long long admitted = S->items - S->phase_items0;
    if (admitted < kMergeLow && S->width < S->max_width)
      S->width = min(2 * S->width, S->max_width);
    else if (admitted > kMergeHigh && S->width > 1)
      S->width >>= 1;
    S->phase_items0 = S->items;
    S->phase += 1;
    S->bucket = S->far_min;
    S->far_min = 0xffffffffu;
    S->far_count[S->far_sel ^ 1] = 0;
  } else {
    S->done = 1;
    return;
  }
  S->steps += 1;
}

// Last kernel of a captured batch: closes its last step, so the state is
// whole when the host reads it.
__global__ void batch_end_kernel(SsspState* S) {
  if (threadIdx.x != 0) return;
  close_step(S);
}

constexpr int kQCap = 2048;

// Holds the output pointer read from the device state: the kernel-argument
// Bufs stays unmodified (a modified copy of it, with its runtime-indexed
// queue arrays, would be kept in scratch memory).
struct RelaxOp {
  Bufs B;
  uint32_t* dist;
  LdsQueue<kQCap>* qn;  // staged near pushes
  LdsQueue<kQCap>* qf;  // staged far pushes
  uint32_t* nearq_out;
  unsigned* near_out_count;
  uint32_t* farq;
  unsigned* far_count;
  uint32_t bucket_end, round, phase, hub_gen;  // bucket_end: last bucket of the phase
  uint32_t far_min = 0xffffffffu;  // this lane's smallest far bucket improved into
  uint32_t hub_min = 0xffffffffu;  // this lane's smallest distance of a hub it listed
  uint32_t near_min = 0xffffffffu;  // this lane's smallest distance improved into the near range
  unsigned long long edges = 0, items = 0;
  // Stale-item filter (the reference's test at SSSP.cpp:138): a near item is
  // expanded only if its distance changed since it was last admitted.  A
  // hub passes into the pending list once per list generation (its edges
  // are counted when the list is expanded).
  __device__ __forceinline__ bool admit(uint32_t u) {
    const uint32_t d = dist[u];
    if (B.done[u] == d) return false;
    B.done[u] = d;
    const uint32_t deg = static_cast<uint32_t>(B.row[u + 1] - B.row[u]);
    if (deg > B.hub) {
      if (!B.defer) return true;
      if (atomicExch(&B.hubmark[u], hub_gen) == hub_gen) return false;
      hub_min = min(hub_min, d);
      return true;
    }
    edges += deg;
    items += 1;
    return true;
  }
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    bool to_near = false, to_far = false;
    uint32_t v = 0;
    if (valid) {
      v = static_cast<uint32_t>(B.col[e]);
      const uint32_t nd = dist[u] + B.w[e];
      if (nd < dist[v]) {
        const uint32_t old = atomicMin(&dist[v], nd);
        if (nd < old) {
          const uint32_t nb = nd >> B.shift;
          if (nb <= bucket_end) {
            near_min = min(near_min, nd);
            to_near = atomicExch(&B.stamp[v], round) != round;
          } else {
            to_far = atomicExch(&B.farmark[v], phase) != phase;
            far_min = min(far_min, nb);
          }
        }
      }
    }
    lq_push(*qn, to_near, v);
    lq_push(*qf, to_far, v);
  }
  __device__ __forceinline__ void flush(bool force) {
    lq_flush(*qn, nearq_out, near_out_count, B.cap, &B.S->overflow, force);
    lq_flush(*qf, farq, far_count, B.cap, &B.S->overflow, force);
  }
  __device__ __forceinline__ void after_step() { flush(false); }
  // End of a launch: queues flushed, far minimum and filter counts reduced
  // once per wavefront.
  // Workgroup-uniform.  The minima and counts are reduced over the
  // workgroup first, so the state words (one cache line) take one atomic
  // per workgroup rather than one per wavefront.
  __device__ __forceinline__ void finish() {
    flush(true);
    constexpr int kW = kBlock / kWave;
    __shared__ uint32_t s_min[3][kW];
    __shared__ unsigned long long s_sum[2][kW];
    const uint32_t fm = wave_min(far_min), hm = wave_min(hub_min), nm = wave_min(near_min);
    const unsigned long long ed = wave_sum(edges), it = wave_sum(items);
    const unsigned w = threadIdx.x / kWave;
    if (lane_id() == 0) {
      s_min[0][w] = fm;
      s_min[1][w] = hm;
      s_min[2][w] = nm;
      s_sum[0][w] = ed;
      s_sum[1][w] = it;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t f = 0xffffffffu, h = 0xffffffffu, nn = 0xffffffffu;
      unsigned long long e = 0, t = 0;
      for (int i = 0; i < kW; ++i) {
        f = min(f, s_min[0][i]);
        h = min(h, s_min[1][i]);
        nn = min(nn, s_min[2][i]);
        e += s_sum[0][i];
        t += s_sum[1][i];
      }
      if (f != 0xffffffffu) atomicMin(&B.S->far_min, f);
      if (h != 0xffffffffu) atomicMin(&B.S->hub_min, h);
      if (nn != 0xffffffffu && B.hub_span) atomicMin(&B.S->near_min_next, nn);  // read only with hub slices
      if (e) atomicAdd(&B.S->edges, e);  // hub steps count edges in every workgroup, items in one
      if (t) atomicAdd(&B.S->items, t);
    }
  }
};

__device__ __forceinline__ RelaxOp make_relax(const Bufs& B, LdsQueue<kQCap>* qn, LdsQueue<kQCap>* qf) {
  SsspState* S = B.S;
  const unsigned q = S->q, fs = S->far_sel;
  return RelaxOp{B, S->dist_out, qn, qf, B.nearq[q ^ 1], &S->near_count[q ^ 1], B.farq[fs], &S->far_count[fs],
                 S->bucket + (S->width - 1), S->round, S->phase, S->hub_gen};
}

// Relax operation over hub edges: counts every valid edge it relaxes.
template <class Op>
struct HubCount {
  Op& op;
  __device__ __forceinline__ void operator()(bool valid, uint32_t u, int32_t e) {
    op.edges += valid ? 1ull : 0ull;
    op(valid, u, e);
  }
  __device__ __forceinline__ void after_step() { op.after_step(); }
};

// Relax step: expand the near items that pass the stale-item filter.
__device__ __forceinline__ void relax_step(const Bufs& B, LdsQueue<kQCap>& qn, LdsQueue<kQCap>& qf) {
  SsspState* S = B.S;
  RelaxOp op = make_relax(B, &qn, &qf);
  const unsigned q = S->q;
  expand_queue<kBlock, RelaxOp, true>(B.row, B.nearq[q], S->near_count[q], B.hub, B.hubq[S->hq], &S->hub_count, B.cap,
                                      &S->overflow, op);
  op.finish();
}

// The hubs of a relax step (hub_count), or with deferral the pending hubs
// of a hub step (hub_expand, 0 outside hub steps), spread over the grid;
// hub_offsets ran over the same count.
__global__ void __launch_bounds__(kBlock) hubs_kernel(Bufs B) {
  SsspState* S = B.S;
  const unsigned hubs = B.defer ? (S->hub_step ? S->hub_expand : 0u) : (S->mode == kRelax ? S->hub_count : 0u);
  if (hubs == 0) return;
  __shared__ LdsQueue<kQCap> qn, qf;
  lq_init(qn);
  lq_init(qf);
  RelaxOp op = make_relax(B, &qn, &qf);
  HubCount<RelaxOp> hc{op};
  expand_hubs<kBlock>(B.row, B.defer ? B.hubx : B.hubq[0], hubs, B.cap, B.hub_off, hc);
  if (blockIdx.x == 0 && threadIdx.x == 0) op.items += min(hubs, B.cap);
  op.finish();
}

// Advance step: far items of the new bucket go to the near queue, later
// ones are kept (one copy per vertex, their buckets fold into the running far
// minimum), settled ones are dropped.
__device__ __forceinline__ void far_split_step(const Bufs& B, LdsQueue<kQCap>& qn, LdsQueue<kQCap>& qk) {
  SsspState* S = B.S;
  const uint32_t* __restrict__ dist = S->dist_out;
  const uint32_t next = S->bucket, last = S->bucket + (S->width - 1);
  const unsigned fs = S->far_sel, q = S->q;
  const unsigned cnt = S->far_count[fs];
  const uint32_t* farq = B.farq[fs];
  const uint32_t round = S->round, phase = S->phase;
  uint32_t kept_min = 0xffffffffu, near_min = 0xffffffffu;
  for (unsigned b0 = blockIdx.x * kBlock; b0 < cnt; b0 += gridDim.x * kBlock) {
    const unsigned i = b0 + threadIdx.x;
    bool to_near = false, keep = false;
    uint32_t v = 0;
    if (i < cnt) {
      v = farq[i];
      const uint32_t dv = dist[v], b = dv >> B.shift;
      if (b >= next && b <= last) {
        near_min = min(near_min, dv);
        to_near = atomicExch(&B.stamp[v], round) != round;
      } else if (b > last) {
        keep = atomicExch(&B.farmark[v], phase) != phase;
        kept_min = min(kept_min, b);
      }
    }
    lq_push(qn, to_near, v);
    lq_push(qk, keep, v);
    lq_flush(qn, B.nearq[q ^ 1], &S->near_count[q ^ 1], B.cap, &S->overflow, false);
    lq_flush(qk, B.farq[fs ^ 1], &S->far_count[fs ^ 1], B.cap, &S->overflow, false);
  }
  lq_flush(qn, B.nearq[q ^ 1], &S->near_count[q ^ 1], B.cap, &S->overflow, true);
  lq_flush(qk, B.farq[fs ^ 1], &S->far_count[fs ^ 1], B.cap, &S->overflow, true);
  kept_min = wave_min(kept_min);
  near_min = wave_min(near_min);
  if (lane_id() == 0 && kept_min != 0xffffffffu) atomicMin(&S->far_min, kept_min);
  if (lane_id() == 0 && near_min != 0xffffffffu && B.hub_span) atomicMin(&S->near_min_next, near_min);
}

// Hub step, first part: the pending hubs at or below hub_limit go to the
// expand list; the others move to the other pending list under the next
// generation's mark (the generation advances when the step closes), so the
// admission filter does not list them twice, and their smallest distance
// becomes the next step's hub_min.
// The two lists are staged in LDS and moved out with one global atomic per
// flush, not one per wavefront: the pending list can hold millions of hubs,
// and every selecting step passes over it.
__global__ void __launch_bounds__(kBlock) hub_select_kernel(Bufs B) {
  SsspState* S = B.S;
  if (!S->hub_step) return;
  __shared__ LdsQueue<kQCap> qs, qk;
  lq_init(qs);
  lq_init(qk);
  const uint32_t* __restrict__ dist = S->dist_out;
  const unsigned cnt = min(S->hub_count, B.cap), hq = S->hq, gen = S->hub_gen + 1;
  const uint32_t limit = S->hub_limit;
  const uint32_t* pending = B.hubq[hq];
  uint32_t kept_min = 0xffffffffu;
  for (unsigned b0 = blockIdx.x * kBlock; b0 < cnt; b0 += gridDim.x * kBlock) {  // workgroup-uniform
    const unsigned i = b0 + threadIdx.x;
    bool sel = false, keep = false;
    uint32_t h = 0;
    if (i < cnt) {
      h = pending[i];
      const uint32_t d = dist[h];
      sel = d <= limit;
      keep = !sel;
      if (keep) {
        B.hubmark[h] = gen;
        kept_min = min(kept_min, d);
      }
    }
    lq_push(qs, sel, h);
    lq_push(qk, keep, h);
    lq_flush(qs, B.hubx, &S->hub_expand, B.cap, &S->overflow, false);
    lq_flush(qk, B.hubq[hq ^ 1], &S->hub_kept, B.cap, &S->overflow, false);
  }
  lq_flush(qs, B.hubx, &S->hub_expand, B.cap, &S->overflow, true);
  lq_flush(qk, B.hubq[hq ^ 1], &S->hub_kept, B.cap, &S->overflow, true);
  kept_min = wave_min(kept_min);
  if (lane_id() == 0 && kept_min != 0xffffffffu) atomicMin(&S->hub_min_next, kept_min);
}

// The work of a step: a relax or an advance (the step's mode picks one; the
// two never share a step, so one launch serves both).
__global__ void __launch_bounds__(kBlock) work_kernel(Bufs B) {
  const unsigned mode = B.S->mode;
  if (mode != kRelax && mode != kAdvance) return;
  __shared__ LdsQueue<kQCap> q0, q1;
  lq_init(q0);
  lq_init(q1);
  if (mode == kRelax)
    relax_step(B, q0, q1);
  else
    far_split_step(B, q0, q1);
}

// `hubs` false: no vertex exceeds the hub threshold, so the hub list stays
// empty and its two launches are left out of the step.
void enqueue_step(const Bufs& B, bool hubs, hipStream_t s) {
  step_kernel<<<1, 64, 0, s>>>(B.S, B.defer, B.hub_span);
  work_kernel<<<hubs ? kGrid : kGridNoHubs, kBlock, 0, s>>>(B);
  if (hubs) {
    if (B.defer) hub_select_kernel<<<kHubSelectGrid, kBlock, 0, s>>>(B);
    launch_hub_offsets<kBlock>(B.row, B.defer ? B.hubx : B.hubq[0], B.defer ? &B.S->hub_expand : &B.S->hub_count,
                               B.cap, B.hub_off, s);
    hubs_kernel<<<kGrid, kBlock, 0, s>>>(B);
  }
  GX_LAUNCHED("sssp step");
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
  Bufs B;
  B.row = g->row;
  B.col = g->col;
  B.w = static_cast<const uint32_t*>(g->w);
  B.dist = dist;
  B.nearq[0] = g->scratch<uint32_t>(WS_Q0, cap);
  B.nearq[1] = g->scratch<uint32_t>(WS_Q1, cap);
  B.farq[0] = g->scratch<uint32_t>(WS_Q2, cap);
  B.farq[1] = g->scratch<uint32_t>(WS_Q3, cap);
  B.hubq[0] = g->scratch<uint32_t>(WS_V0, cap);
  B.hubq[1] = g->scratch<uint32_t>(WS_BITMAP0, cap);
  B.hubx = g->scratch<uint32_t>(WS_BITMAP1, cap);
  B.stamp = g->scratch<uint32_t>(WS_V1, cap);
  B.farmark = g->scratch<uint32_t>(WS_V2, cap);
  B.done = g->scratch<uint32_t>(WS_V3, cap);
  B.hubmark = g->scratch<uint32_t>(WS_V4, cap);
  B.S = g->scratch<SsspState>(WS_COUNTERS, 1);
  B.shift = shift;
  B.cap = cap;
  B.hub_off = g->scratch<uint32_t>(WS_HUBOFF, hub_off_words(cap));
  B.hub = hub_degree();
  const int64_t refine = tune_effective("sssp.hub_refine", n, g->m);
  B.defer = static_cast<uint32_t>(tune_effective("sssp.defer_hubs", n, g->m)) | (refine > 0 ? 1u : 0u);
  B.hub_span = refine > 0 ? 1u << std::min<int64_t>(refine - 1, 31) : 0u;
  const bool hubs = max_degree(g) > static_cast<int64_t>(B.hub);
  // Graphs with hubs run on the degree-ordered copy (frontier.relabel).
  const bool relabel = hubs && tune_effective("frontier.relabel", n, g->m) != 0;
  uint32_t run_src = src;
  if (relabel) {
    build_relabeled(g);
    B.row = g->relab.row.as<int32_t>();
    B.col = g->relab.col.as<int32_t>();
    B.w = g->relab.w.as<uint32_t>();
    B.dist = g->scratch<uint32_t>(WS_V5, cap);
    run_src = read_scalar(g, g->relab.new_of_old.as<uint32_t>() + src);
  }
  SsspState* host = static_cast<SsspState*>(g->host_pinned(sizeof(SsspState)));

  RegionTimer region(s);
  KernelTimer ktimer(s);
  const unsigned max_width = static_cast<unsigned>(tune_effective("sssp.max_merge", n, g->m));
  sssp_init_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(B.dist, B.stamp, B.farmark, B.done, B.hubmark, n,
                                                                 run_src);
  GX_LAUNCHED("sssp_init_kernel");
  sssp_reset_kernel<<<1, 64, 0, s>>>(B.S, B.nearq[0], run_src, B.dist, max_width);
  GX_LAUNCHED("sssp_reset_kernel");
  uint64_t key = 0;
  for (const void* p : {(const void*)B.row, (const void*)B.col, (const void*)B.w,
                        (const void*)B.nearq[0], (const void*)B.nearq[1], (const void*)B.farq[0],
                        (const void*)B.farq[1], (const void*)B.hubq[0], (const void*)B.hubq[1], (const void*)B.hubx,
                        (const void*)B.stamp, (const void*)B.farmark,
                        (const void*)B.done, (const void*)B.hubmark, (const void*)B.S, (const void*)B.hub_off})
    key = mix_key(key, reinterpret_cast<uintptr_t>(p));
  key = mix_key(mix_key(mix_key(mix_key(mix_key(mix_key(key, shift), cap), B.hub), hubs ? 1u : 0u), B.defer),
                B.hub_span);
  uint32_t replays = 0;
  for (;;) {
    ktimer.begin();
    g->replay(g->sssp_batch, key, [&](hipStream_t cs) {
      for (int i = 0; i < kStepsPerGraph; ++i) enqueue_step(B, hubs, cs);
      batch_end_kernel<<<1, 64, 0, cs>>>(B.S);
      GX_LAUNCHED("batch_end_kernel");
    });
    ktimer.end();
    ++replays;
    GX_HIP(hipMemcpyAsync(host, B.S, sizeof(SsspState), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    if (host->overflow) fail(GX_ERR_CAPACITY, "SSSP queue overflow");
    if (host->done) break;
  }
  if (relabel) unpermute_u32(B.dist, g->relab.new_of_old.as<uint32_t>(), dist, n, s);
  double ms = region.stop();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->iterations = host->steps;
    st->launches = 2 + replays * (kStepsPerGraph * (hubs ? (B.defer ? 5 : 4) : 2) + 1);
    st->edges_visited = host->edges;
    st->items = host->items;
    st->elapsed_ms = ms;
    st->main_kernel_ms = ktimer.total_ms();
    st->main_kernel_launches = ktimer.count();
  }
  GX_API_END
}
