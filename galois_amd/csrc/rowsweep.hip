// Plan tables of the row sweep (rowsweep.hpp): window first rows, long rows
// and their pieces.  Built once per row-offset array and cached in the graph.
#include <algorithm>

#include "rowsweep.hpp"

namespace gx {
namespace rs {
namespace {

__global__ void win_rows_kernel(const int32_t* __restrict__ rowoff, int64_t n, int64_t nwin,
                                int32_t* __restrict__ win_rows) {
  int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; k <= nwin; k += stride) {
    if (k == nwin) {
      win_rows[k] = static_cast<int32_t>(n);
      continue;
    }
    const int64_t target = k * kWin;
    int64_t lo = 0, hi = n;  // first r in [0,n) with rowoff[r] >= target, else n
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (rowoff[mid] >= target) hi = mid; else lo = mid + 1;
    }
    win_rows[k] = static_cast<int32_t>(lo);
  }
}

__global__ void long_flags_kernel(const int32_t* __restrict__ rowoff, int64_t n, uint32_t* __restrict__ flag) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) flag[v] = (rowoff[v + 1] - rowoff[v] > kWin) ? 1u : 0u;
}

__global__ void long_compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int64_t n,
                                    uint32_t* __restrict__ out) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride)
    if (flag[v]) out[pos[v]] = static_cast<uint32_t>(v);
}

__global__ void piece_count_kernel(const int32_t* __restrict__ rowoff, const uint32_t* __restrict__ long_rows,
                                   int64_t nlong, uint32_t* __restrict__ cnt) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < nlong) {
    const uint32_t r = long_rows[i];
    const int64_t d = rowoff[r + 1] - rowoff[r];
    cnt[i] = static_cast<uint32_t>((d + kPiece - 1) / kPiece);
  } else if (i == nlong) {
    cnt[i] = 0;
  }
}

__global__ void piece_row_kernel(const uint32_t* __restrict__ piece_first, int64_t nlong,
                                 uint32_t* __restrict__ piece_row) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < nlong; i += stride)
    for (uint32_t p = piece_first[i]; p < piece_first[i + 1]; ++p) piece_row[p] = static_cast<uint32_t>(i);
}

}  // namespace

// Table layout in cache.buf: win_rows[nwin+1] | long_rows[max_long] |
// piece_first[max_long+1] | piece_row[max_pieces].
RowPlan build_plan(gx_graph* g, const int32_t* rowoff, const int32_t* col, int64_t n, int64_t m, PullPlan& cache) {
  const int64_t nwin = m / kWin + 1;
  const int64_t max_long = m / kWin + 1;
  const int64_t max_pieces = m / kPiece + max_long + 1;
  auto view = [&]() {
    RowPlan P{};
    int32_t* base = cache.buf.as<int32_t>();
    P.rowoff = rowoff;
    P.col = col;
    P.len = m;
    P.nwin = nwin;
    P.win_rows = base;
    P.long_rows = reinterpret_cast<uint32_t*>(base + (nwin + 1));
    P.piece_first = P.long_rows + max_long;
    P.piece_row = P.piece_first + (max_long + 1);
    P.nlong = cache.nlong;
    P.npieces = cache.npieces;
    return P;
  };
  if (cache.ready) return view();
  GX_REQUIRE(n > 0, "row sweep over an empty graph");
  hipStream_t s = g->stream;
  cache.buf.ensure(((nwin + 1) + max_long + (max_long + 1) + max_pieces) * sizeof(int32_t));
  cache.nlong = cache.npieces = 0;
  RowPlan P = view();
  win_rows_kernel<<<grid_for(nwin + 1, kBlock, 4096), kBlock, 0, s>>>(rowoff, n, nwin, P.win_rows);
  GX_LAUNCHED("win_rows_kernel");
  DevBuf flag, pos;
  flag.ensure(n * 4);
  pos.ensure(n * 4);
  long_flags_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(rowoff, n, flag.as<uint32_t>());
  GX_LAUNCHED("long_flags_kernel");
  exclusive_scan_u32(flag.as<uint32_t>(), pos.as<uint32_t>(), n, s);
  uint32_t lastpos = 0, lastflag = 0;
  read_bytes(g, &lastpos, pos.as<uint32_t>() + n - 1, 4);
  read_bytes(g, &lastflag, flag.as<uint32_t>() + n - 1, 4);
  const int64_t nlong = static_cast<int64_t>(lastpos) + lastflag;
  GX_REQUIRE(nlong <= max_long, "row sweep long-row table bound exceeded");
  int64_t npieces = 0;
  if (nlong > 0) {
    long_compact_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(flag.as<uint32_t>(), pos.as<uint32_t>(), n,
                                                                      P.long_rows);
    GX_LAUNCHED("long_compact_kernel");
    DevBuf cnt;
    cnt.ensure((nlong + 1) * 4);
    piece_count_kernel<<<grid_for(nlong + 1, kBlock), kBlock, 0, s>>>(rowoff, P.long_rows, nlong,
                                                                      cnt.as<uint32_t>());
    GX_LAUNCHED("piece_count_kernel");
    exclusive_scan_u32(cnt.as<uint32_t>(), P.piece_first, nlong + 1, s);
    uint32_t np = 0;
    read_bytes(g, &np, P.piece_first + nlong, 4);
    npieces = np;
    GX_REQUIRE(npieces <= max_pieces, "row sweep piece table bound exceeded");
    piece_row_kernel<<<grid_for(nlong, kBlock, 4096), kBlock, 0, s>>>(P.piece_first, nlong, P.piece_row);
    GX_LAUNCHED("piece_row_kernel");
  }
  GX_HIP(hipStreamSynchronize(s));
  cache.nlong = nlong;
  cache.npieces = npieces;
  cache.ready = true;
  return view();
}

}  // namespace rs
}  // namespace gx
