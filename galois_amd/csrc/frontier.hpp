// Load-balanced neighbour expansion shared by the frontier-driven kernels
// (BFS, SSSP, k-core, CC hooking, BC forward).
//
// A workgroup of BLOCK lanes takes BLOCK frontier vertices, exclusive-scans
// their degrees in LDS and walks the concatenated edge list with all lanes;
// the owning vertex of edge slot e is found by binary search over the LDS
// offsets.  Vertices whose degree exceeds `hub_thresh` are not expanded here:
// they are appended to a hub queue and their edges are later spread over the
// whole grid by expand_hubs.  This gives three effective tiers — short
// adjacencies packed many-per-wavefront, medium ones spread over the
// workgroup, hubs spread over the grid.
//
// The per-edge operation is called by every lane of the workgroup the same
// number of times (with `valid` false for padding lanes), so it may use
// wavefront-collective operations (ballot, shuffles).  After every step of
// BLOCK edges, every lane calls op.after_step() (workgroup-uniform), where an
// operation may flush its LDS-staged output (LdsQueue below).
#pragma once

#include "gx_device.hpp"

namespace gx {

// Identity frontier (q null): vertex first + i for i in [0, count).
// Otherwise q[i].  With ADMIT, op.admit(v) is asked once per frontier vertex
// and a vertex it rejects contributes no edges (a fused filter).
template <int BLOCK, class Op, bool ADMIT = false>
__device__ __forceinline__ void expand_queue(const int32_t* __restrict__ row, const uint32_t* __restrict__ q,
                                             unsigned qcount, uint32_t hub_thresh, uint32_t* __restrict__ hubq,
                                             unsigned* hub_count, unsigned cap, unsigned* overflow, Op& op,
                                             uint32_t first = 0) {
  __shared__ int32_t s_start[BLOCK];
  __shared__ uint32_t s_off[BLOCK];
  __shared__ uint32_t s_src[BLOCK];
  __shared__ uint64_t s_scan[BLOCK / kWave + 1];
  __shared__ unsigned s_hub_base;
  const unsigned tid = threadIdx.x;
  for (unsigned base = blockIdx.x * BLOCK; base < qcount; base += gridDim.x * BLOCK) {
    const unsigned i = base + tid;
    uint32_t v = 0, deg = 0;
    int32_t start = 0;
    bool hub = false;
    if (i < qcount) {
      v = q ? q[i] : first + i;
      bool in = true;
      if constexpr (ADMIT) in = op.admit(v);
      if (in) {
        start = row[v];
        deg = static_cast<uint32_t>(row[v + 1] - start);
        if (deg > hub_thresh) {
          hub = true;
          deg = 0;
        }
      }
    }
    // One scan gives both the edge offsets (low 40 bits) and the hub slots
    // (high bits), so the hub list takes one global atomic per workgroup
    // batch instead of one per wavefront.
    uint64_t total64;
    const uint64_t ex64 = block_exclusive_scan<BLOCK>(static_cast<uint64_t>(deg) | (static_cast<uint64_t>(hub) << 40),
                                                      s_scan, &total64);
    constexpr uint64_t kLow = (uint64_t(1) << 40) - 1;
    const uint32_t excl = static_cast<uint32_t>(ex64 & kLow), total = static_cast<uint32_t>(total64 & kLow);
    const uint32_t hub_total = static_cast<uint32_t>(total64 >> 40);  // workgroup-uniform
    if (hubq && hub_total && tid == 0) s_hub_base = atomicAdd(hub_count, hub_total);
    s_start[tid] = start;
    s_off[tid] = excl;
    s_src[tid] = v;
    __syncthreads();
    if (hubq && hub) {
      const unsigned slot = s_hub_base + static_cast<uint32_t>(ex64 >> 40);
      if (slot < cap)
        hubq[slot] = v;
      else
        atomicOr(overflow, 1u);
    }
    for (uint32_t e0 = 0; e0 < total; e0 += BLOCK) {
      const uint32_t e = e0 + tid;
      const bool valid = e < total;
      uint32_t u = 0;
      int32_t edge = 0;
      if (valid) {
        int lo = 0, hi = BLOCK - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_off[mid] <= e) lo = mid; else hi = mid - 1;
        }
        u = s_src[lo];
        edge = s_start[lo] + static_cast<int32_t>(e - s_off[lo]);
      }
      op(valid, u, edge);
      op.after_step();
    }
    __syncthreads();
  }
}

// Step offsets of the hub list for expand_hubs, computed by
// kHubSlices workgroups: the hub list hubq[0..hubs) (hubs = min(*hub_count,
// cap); the count runs past cap when the hub queue overflowed, which the
// caller reports) is cut into kHubSlices slices of 2^hub_chunk_shift(hubs)
// hubs (the last ones may be empty), and slice k
// writes off[h] = sum over hubs h' < h of its own slice of ceil(deg(hubq[h'])
// / BLOCK) for its hubs h, and its total to off[cap + 1 + k].  expand_hubs
// adds the slice bases; it must be given the same `cap`.  `off` holds
// hub_off_words(cap) words.
constexpr int kHubSlices = 256;
constexpr int64_t hub_off_words(int64_t cap) { return cap + 1 + kHubSlices; }

// log2 of the slice length: the smallest power of two with
// kHubSlices << shift >= hubs (a shift, not a division, in the search).
__device__ __forceinline__ unsigned hub_chunk_shift(unsigned hubs) {
  unsigned sh = 0;
  while ((static_cast<uint64_t>(kHubSlices) << sh) < hubs) ++sh;
  return sh;
}

constexpr int kHubScanThreads = 256;

template <int BLOCK>
__global__ void __launch_bounds__(kHubScanThreads) hub_offsets_kernel(const int32_t* __restrict__ row,
                                                                      const uint32_t* __restrict__ hubq,
                                                                      const unsigned* __restrict__ hub_count,
                                                                      unsigned cap, uint32_t* __restrict__ off) {
  constexpr int NT = kHubScanThreads;
  __shared__ uint32_t lds[NT / kWave + 1];
  const unsigned hubs = min(*hub_count, cap);
  const unsigned sh = hub_chunk_shift(hubs);
  const unsigned b0 = static_cast<unsigned>(min(static_cast<uint64_t>(blockIdx.x) << sh, static_cast<uint64_t>(hubs)));
  const unsigned b1 = static_cast<unsigned>(min(static_cast<uint64_t>(b0) + (uint64_t(1) << sh), static_cast<uint64_t>(hubs)));
  uint32_t carry = 0;
  for (unsigned base = b0; base < b1; base += NT) {
    const unsigned i = base + threadIdx.x;
    uint32_t c = 0;
    if (i < b1) {
      const uint32_t v = hubq[i];
      c = static_cast<uint32_t>((row[v + 1] - row[v] + BLOCK - 1) / BLOCK);
    }
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<NT>(c, lds, &total);
    if (i < b1) off[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) off[cap + 1 + blockIdx.x] = carry;
}

template <int BLOCK>
inline void launch_hub_offsets(const int32_t* row, const uint32_t* hubq, const unsigned* hub_count, unsigned cap,
                               uint32_t* off, hipStream_t s) {
  hub_offsets_kernel<BLOCK><<<kHubSlices, kHubScanThreads, 0, s>>>(row, hubq, hub_count, cap, off);
}

// Edges of the hub vertices hubq[0..hubs), cut into BLOCK-edge steps (a
// hub's steps are consecutive).  Each workgroup takes a contiguous range of
// steps: one binary search over the hub offsets (those of hub_offsets_kernel
// plus the scanned slice totals) finds the hub of its first step, and the
// following steps advance to the next hub when they pass its first step, so
// a step costs no search.
template <int BLOCK, class Op>
__device__ __forceinline__ void expand_hubs(const int32_t* __restrict__ row, const uint32_t* __restrict__ hubq,
                                            unsigned hub_count, unsigned cap, const uint32_t* __restrict__ off,
                                            Op& op) {
  static_assert(BLOCK >= kHubSlices, "one slice total per lane");
  __shared__ uint32_t s_base[kHubSlices];
  __shared__ uint32_t s_scan[BLOCK / kWave + 1];
  __shared__ uint32_t s_total;
  const unsigned hubs = min(hub_count, cap);
  if (hubs == 0) return;
  const unsigned sh = hub_chunk_shift(hubs);
  {
    const uint32_t t = threadIdx.x < kHubSlices ? off[cap + 1 + threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<BLOCK>(t, s_scan, &total);
    if (threadIdx.x < kHubSlices) s_base[threadIdx.x] = ex;
    if (threadIdx.x == 0) s_total = total;
    __syncthreads();
  }
  const uint32_t total = s_total;
  const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
  const uint32_t c0 = static_cast<uint32_t>(min(static_cast<uint64_t>(blockIdx.x) * per, static_cast<uint64_t>(total)));
  const uint32_t c1 = static_cast<uint32_t>(min(static_cast<uint64_t>(c0) + per, static_cast<uint64_t>(total)));
  if (c0 >= c1) return;
  auto first_step = [&](unsigned h) { return off[h] + s_base[h >> sh]; };
  unsigned lo = 0, hi = hubs - 1;  // last h with first_step(h) <= c0 (steps >= 1 per hub: increasing)
  while (lo < hi) {
    const unsigned mid = (lo + hi + 1) >> 1;
    if (first_step(mid) <= c0) lo = mid; else hi = mid - 1;
  }
  unsigned h = lo;
  uint32_t first = first_step(h);
  uint32_t next = h + 1 < hubs ? first_step(h + 1) : total;
  uint32_t v = hubq[h];
  int64_t rb = row[v], re = row[v + 1];
  for (uint32_t c = c0; c < c1; ++c) {
    if (c >= next) {  // workgroup-uniform; the next hub has >= 1 step
      ++h;
      first = next;
      next = h + 1 < hubs ? first_step(h + 1) : total;
      v = hubq[h];
      rb = row[v];
      re = row[v + 1];
    }
    const int64_t e = rb + static_cast<int64_t>(c - first) * BLOCK + threadIdx.x;
    const bool valid = e < re;
    op(valid, v, valid ? static_cast<int32_t>(e) : 0);
    op.after_step();
  }
}

// Workgroup-private output queue in LDS.  Lanes append with one LDS atomic
// per wavefront; the workgroup moves the items to the global queue with one
// global atomic per flush (coalesced stores), instead of one global atomic
// per wavefront per step on a single hot counter.
template <int CAP>
struct LdsQueue {
  uint32_t items[CAP];
  unsigned count;
  unsigned gbase;
};

template <int CAP>
__device__ __forceinline__ void lq_init(LdsQueue<CAP>& q) {
  if (threadIdx.x == 0) q.count = 0;
  __syncthreads();
}

// Room is guaranteed by flushing whenever fewer than one step's worth
// (blockDim.x items) of space is left.
template <int CAP>
__device__ __forceinline__ void lq_push(LdsQueue<CAP>& q, bool pred, uint32_t v) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const unsigned leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  unsigned base = 0;
  if (lane_id() == leader) base = atomicAdd(&q.count, static_cast<unsigned>(__popcll(mask)));
  base = __shfl(base, leader);
  if (pred) q.items[base + lanes_below(mask)] = v;
}

// Workgroup-uniform: every lane of the workgroup must call it.  Flushes when
// `force` or when the queue is nearly full.
template <int CAP>
__device__ __forceinline__ void lq_flush(LdsQueue<CAP>& q, uint32_t* __restrict__ gq, unsigned* gcount, unsigned cap,
                                         unsigned* overflow, bool force) {
  __syncthreads();
  const unsigned cnt = q.count;
  __syncthreads();
  if (!force && cnt + blockDim.x <= CAP) return;
  if (cnt == 0) return;
  if (threadIdx.x == 0) q.gbase = atomicAdd(gcount, cnt);
  __syncthreads();
  const unsigned base = q.gbase;
  for (unsigned i = threadIdx.x; i < cnt; i += blockDim.x) {
    if (base + i < cap)
      gq[base + i] = q.items[i];
    else
      atomicOr(overflow, 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) q.count = 0;
  __syncthreads();
}

// Segmented inclusive sum across a wavefront where equal `key`s form
// contiguous runs of lanes (invalid lanes carry a key no valid lane has).
// Returns the run total in the last lane of each run (is_tail set).
template <class T>
__device__ __forceinline__ T wave_segmented_sum(uint32_t key, T v, bool& is_tail) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const T ov = __shfl_up(v, off);
    const uint32_t ok = __shfl_up(key, off);
    if (lane >= static_cast<unsigned>(off) && ok == key) v += ov;
  }
  const uint32_t nk = __shfl_down(key, 1);
  is_tail = (lane == kWave - 1) || (nk != key);
  return v;
}

}  // namespace gx
