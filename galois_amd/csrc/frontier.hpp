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
// wavefront-collective operations (ballot, shuffles).
#pragma once

#include "gx_device.hpp"

namespace gx {

// Identity frontier: vertex i of [0, count).  Otherwise q[i].
template <int BLOCK, class Op>
__device__ __forceinline__ void expand_queue(const int32_t* __restrict__ row, const uint32_t* __restrict__ q,
                                             unsigned qcount, uint32_t hub_thresh, uint32_t* __restrict__ hubq,
                                             unsigned* hub_count, unsigned cap, unsigned* overflow, Op& op) {
  __shared__ int32_t s_start[BLOCK];
  __shared__ uint32_t s_off[BLOCK];
  __shared__ uint32_t s_src[BLOCK];
  __shared__ uint32_t s_scan[BLOCK / kWave + 1];
  const unsigned tid = threadIdx.x;
  for (unsigned base = blockIdx.x * BLOCK; base < qcount; base += gridDim.x * BLOCK) {
    const unsigned i = base + tid;
    uint32_t v = 0, deg = 0;
    int32_t start = 0;
    bool hub = false;
    if (i < qcount) {
      v = q ? q[i] : i;
      start = row[v];
      deg = static_cast<uint32_t>(row[v + 1] - start);
      if (deg > hub_thresh) {
        hub = true;
        deg = 0;
      }
    }
    if (hubq) wave_append<uint32_t>(hub, v, hubq, hub_count, cap, overflow);
    uint32_t total;
    const uint32_t excl = block_exclusive_scan<BLOCK>(deg, s_scan, &total);
    s_start[tid] = start;
    s_off[tid] = excl;
    s_src[tid] = v;
    __syncthreads();
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
    }
    __syncthreads();
  }
}

// Edges of the hub vertices hubq[0..hubs), spread over every workgroup.
template <int BLOCK, class Op>
__device__ __forceinline__ void expand_hubs(const int32_t* __restrict__ row, const uint32_t* __restrict__ hubq,
                                            unsigned hubs, Op& op) {
  for (unsigned h = 0; h < hubs; ++h) {
    const uint32_t v = hubq[h];
    const int32_t start = row[v], end = row[v + 1];
    for (int64_t e0 = start + static_cast<int64_t>(blockIdx.x) * BLOCK; e0 < end;
         e0 += static_cast<int64_t>(gridDim.x) * BLOCK) {
      const int64_t e = e0 + threadIdx.x;
      const bool valid = e < end;
      op(valid, v, valid ? static_cast<int32_t>(e) : 0);
    }
  }
}

}  // namespace gx
