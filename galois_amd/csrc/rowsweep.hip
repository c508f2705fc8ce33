// Plan table of the row sweep (rowsweep.hpp): the merge-path coordinates
// (first row, first edge) of every tile.  Built once per row-offset array and
// cached in the graph.
#include <algorithm>

#include "rowsweep.hpp"

namespace gx {
namespace rs {
namespace {

// Item d of the merge path (row ends and edges of a sequential CSR walk, a
// row's end after its last edge): rows ended before item d = the largest r
// with r == 0 or rowoff[r] + r <= d; edges before it = d - r.
__global__ void tile_coords_kernel(const int32_t* __restrict__ rowoff, int64_t n, int64_t m, int64_t ntiles,
                                   int32_t* __restrict__ tile_row, int32_t* __restrict__ tile_edge) {
  int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; t <= ntiles; t += stride) {
    const int64_t d = std::min<int64_t>(t * kTile, n + m);
    int64_t lo = std::max<int64_t>(0, d - m), hi = std::min<int64_t>(d, n);
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (static_cast<int64_t>(rowoff[mid]) + mid <= d) lo = mid; else hi = mid - 1;
    }
    tile_row[t] = static_cast<int32_t>(lo);
    tile_edge[t] = static_cast<int32_t>(d - lo);
  }
}

}  // namespace

// Table layout in cache.buf: tile_row[ntiles+1] | tile_edge[ntiles+1].
RowPlan build_plan(gx_graph* g, const int32_t* rowoff, const int32_t* col, int64_t n, int64_t m, PullPlan& cache) {
  GX_REQUIRE(n > 0, "row sweep over an empty graph");
  const int64_t ntiles = (n + m + kTile - 1) / kTile;
  auto view = [&]() {
    RowPlan P{};
    P.rowoff = rowoff;
    P.col = col;
    P.n = n;
    P.len = m;
    P.tile_row = cache.buf.as<int32_t>();
    P.tile_edge = P.tile_row + (ntiles + 1);
    P.ntiles = ntiles;
    return P;
  };
  if (cache.ready) return view();
  cache.buf.ensure(2 * (ntiles + 1) * sizeof(int32_t));
  RowPlan P = view();
  tile_coords_kernel<<<grid_for(ntiles + 1, kBlock, 4096), kBlock, 0, g->stream>>>(rowoff, n, m, ntiles, P.tile_row,
                                                                                  P.tile_edge);
  GX_LAUNCHED("tile_coords_kernel");
  GX_HIP(hipStreamSynchronize(g->stream));
  cache.ntiles = ntiles;
  cache.ready = true;
  return view();
}

}  // namespace rs
}  // namespace gx
