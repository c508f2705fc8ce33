// Device graph lifecycle, CSR/CSC construction and the seeded generators.
//
// Graph construction (sorting edge keys, transposing) is not on the analytics
// hot path; it uses rocPRIM's radix sort and scans.  Everything that runs per
// algorithm step lives in the per-algorithm files and is hand-written.
//
// Reference behaviour kept:
//  - transpose = graph-convert -transpose (tools/graph-convert/
//    graph-convert.cpp:1831-1880): in-edges of v ordered by ascending source;
//  - cleanup = graph-convert -cleanup (graph-convert.cpp:1902-1994): self-loops
//    and duplicate edges removed, adjacency sorted;
//  - symmetrise = FileGraph makeSymmetric (libgalois/include/galois/graphs/
//    FileGraph.h:764-824): the reverse of every edge added, same weight.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>


#include "gx_device.hpp"
#include "gx_internal.hpp"

using namespace gx;

gx_graph::~gx_graph() {
  if (pinned) (void)hipHostFree(pinned);
  if (capture_stream) (void)hipStreamDestroy(capture_stream);
}

hipStream_t gx_graph::capture() {
  if (!capture_stream) GX_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
  return capture_stream;
}

void* gx_graph::host_pinned(size_t bytes) {
  if (bytes > pinned_bytes) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
    size_t b = std::max<size_t>(bytes, 4096);
    GX_HIP(hipHostMalloc(&pinned, b, hipHostMallocDefault));
    pinned_bytes = b;
  }
  return pinned;
}

namespace gx {

void read_bytes(gx_graph* g, void* host, const void* dev, size_t bytes) {
  void* h = g->host_pinned(bytes);
  GX_HIP(hipMemcpyAsync(h, dev, bytes, hipMemcpyDeviceToHost, g->stream));
  GX_HIP(hipStreamSynchronize(g->stream));
  std::memcpy(host, h, bytes);
}

template <class T>
static void fill(T* p, T v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  fill_kernel<T><<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(p, v, n);
  GX_LAUNCHED("fill_kernel");
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s) {
  scan_u32_sum(in, out, count, s);
}

__global__ void max_degree_kernel(const int32_t* __restrict__ row, int64_t n, unsigned* __restrict__ out) {
  unsigned mx = 0;
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x)
    mx = max(mx, static_cast<unsigned>(row[v + 1] - row[v]));
  mx = wave_max(mx);
  if (lane_id() == 0 && mx) atomicMax(out, mx);
}

int64_t max_degree(gx_graph* g) {
  if (g->max_deg >= 0) return g->max_deg;
  unsigned mx = 0;
  if (g->n > 0) {
    DevBuf d;
    d.ensure(sizeof(unsigned));
    GX_HIP(hipMemsetAsync(d.p, 0, sizeof(unsigned), g->stream));
    max_degree_kernel<<<grid_for(g->n, 256, 4096), 256, 0, g->stream>>>(g->row, g->n, d.as<unsigned>());
    GX_LAUNCHED("max_degree_kernel");
    GX_HIP(hipMemcpyAsync(&mx, d.p, sizeof(unsigned), hipMemcpyDeviceToHost, g->stream));
    GX_HIP(hipStreamSynchronize(g->stream));
  }
  g->max_deg = mx;
  return g->max_deg;
}

void sort_keys_u64(const uint64_t* in, uint64_t* out, int64_t count, int end_bit, hipStream_t s) {
  radix_sort_u64(in, out, nullptr, nullptr, count, end_bit, s);
}

static int bits_for(int64_t n) {
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  return b;
}

// ---------------------------------------------------------------- kernels

__global__ void outdeg_kernel(const int32_t* __restrict__ row, uint32_t* __restrict__ deg, int64_t n) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) deg[v] = static_cast<uint32_t>(row[v + 1] - row[v]);
}

// head[row[v]] = v for every vertex with edges (head pre-filled with 0); an
// inclusive max-scan of head then yields the source vertex of every edge.
__global__ void row_heads_kernel(const int32_t* __restrict__ row, uint32_t* __restrict__ head, int64_t n) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride)
    if (row[v + 1] > row[v]) head[row[v]] = static_cast<uint32_t>(v);
}

__global__ void transpose_keys_kernel(const int32_t* __restrict__ col, const uint32_t* __restrict__ src,
                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ ids, int64_t m,
                                      int bits) {
  int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; e < m; e += stride) {
    keys[e] = (static_cast<uint64_t>(static_cast<uint32_t>(col[e])) << bits) | src[e];
    ids[e] = static_cast<uint32_t>(e);
  }
}

// Row offsets from keys sorted by their hi ids (key >> bits < n): the last
// edge of each hi id x writes ends[x] = its index + 1, where ends = row + 1
// starts zeroed; a running maximum over ends then gives every vertex the
// end of the last nonempty vertex at or before it, i.e. row[x + 1].  Work
// and depth do not depend on how long the runs of edgeless vertices are.
__global__ void run_ends_kernel(const uint64_t* __restrict__ keys, int64_t m, int bits, uint32_t* __restrict__ ends) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < m; i += stride) {
    const uint64_t hi = keys[i] >> bits;
    if (i == m - 1 || (keys[i + 1] >> bits) != hi) ends[hi] = static_cast<uint32_t>(i + 1);
  }
}

void offsets_from_sorted(const uint64_t* keys, int64_t m, int bits, int64_t n, int32_t* row, hipStream_t s) {
  fill<int32_t>(row, 0, n + 1, s);
  if (m == 0 || n == 0) return;
  uint32_t* ends = reinterpret_cast<uint32_t*>(row + 1);
  run_ends_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(keys, m, bits, ends);
  GX_LAUNCHED("run_ends_kernel");
  inclusive_max_scan_u32(ends, ends, n, s);
}

__global__ void lo_bits_kernel(const uint64_t* __restrict__ keys, int32_t* __restrict__ out, int64_t m,
                               uint64_t mask) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < m; i += stride) out[i] = static_cast<int32_t>(keys[i] & mask);
}

// keep[i] = 1 when key i is the first of its run and not a self-loop.
__global__ void unique_flags_kernel(const uint64_t* __restrict__ keys, int64_t cnt, int bits,
                                    uint32_t* __restrict__ keep) {
  uint64_t mask = (uint64_t(1) << bits) - 1;
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < cnt; i += stride) {
    uint64_t k = keys[i];
    bool first = (i == 0) || (keys[i - 1] != k);
    bool loop = (k >> bits) == (k & mask);
    keep[i] = (first && !loop) ? 1u : 0u;
  }
}

__global__ void compact_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ keep,
                               const uint32_t* __restrict__ pos, int64_t cnt, uint64_t* __restrict__ out) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < cnt; i += stride)
    if (keep[i]) out[pos[i]] = keys[i];
}

// RMAT endpoint of generated edge i: `scale` quadrant draws, two 32-bit draws
// per 64-bit hash, thresholds pre-scaled to 2^32 so host and device agree.
__host__ __device__ inline void rmat_edge(uint64_t seed, uint64_t i, int scale, uint64_t ta, uint64_t tab,
                                          uint64_t tabc, uint64_t& u, uint64_t& v) {
  u = 0;
  v = 0;
  for (int l = 0; l < scale; l += 2) {
    uint64_t h = hash2(seed, i, static_cast<uint64_t>(l));
    for (int k = 0; k < 2 && l + k < scale; ++k) {
      uint64_t r = (k == 0) ? (h & 0xffffffffull) : (h >> 32);
      uint64_t bit = uint64_t(1) << (scale - 1 - (l + k));
      if (r < ta) {
      } else if (r < tab) {
        v |= bit;
      } else if (r < tabc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
  }
}

__global__ void rmat_keys_kernel(uint64_t* __restrict__ keys, int64_t mgen, int scale, uint64_t ta, uint64_t tab,
                                 uint64_t tabc, uint64_t seed, int symmetric) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < mgen; i += stride) {
    uint64_t u, v;
    rmat_edge(seed, static_cast<uint64_t>(i), scale, ta, tab, tabc, u, v);
    if (symmetric) {
      keys[2 * i] = (u << scale) | v;
      keys[2 * i + 1] = (v << scale) | u;
    } else {
      keys[i] = (u << scale) | v;
    }
  }
}

// Grid lattice edge i (horizontal edges first, then vertical, then
// shortcuts); dropped lattice edges become the self-loop key 0 (removed).
__global__ void grid_keys_kernel(uint64_t* __restrict__ keys, uint32_t side, int64_t nlat, int64_t nshort,
                                 uint64_t drop_thresh, uint64_t seed, int bits) {
  int64_t total = nlat + nshort;
  int64_t hcount = static_cast<int64_t>(side) * (side - 1);
  uint64_t n = static_cast<uint64_t>(side) * side;
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < total; i += stride) {
    uint64_t u, v;
    bool keep = true;
    if (i < nlat) {
      if (i < hcount) {
        uint64_t r = i / (side - 1), c = i % (side - 1);
        u = r * side + c;
        v = u + 1;
      } else {
        uint64_t j = i - hcount;
        uint64_t r = j / side, c = j % side;
        u = r * side + c;
        v = u + side;
      }
      keep = (hash2(seed, 1, static_cast<uint64_t>(i)) >> 32) >= drop_thresh;
    } else {
      uint64_t h = hash2(seed, 2, static_cast<uint64_t>(i - nlat));
      u = ((h & 0xffffffffull) * n) >> 32;
      v = ((h >> 32) * n) >> 32;
    }
    if (!keep) { u = 0; v = 0; }
    keys[2 * i] = (u << bits) | v;
    keys[2 * i + 1] = (v << bits) | u;
  }
}

// Weight of an undirected pair: uniform integer in [lo, hi].
__host__ __device__ inline uint32_t pair_weight(uint64_t seed, uint64_t u, uint64_t v, uint32_t lo, uint32_t hi) {
  uint64_t a = u < v ? u : v, b = u < v ? v : u;
  uint64_t h = hash2(seed ^ 0x5757575757575757ull, a, b) >> 32;
  uint64_t span = static_cast<uint64_t>(hi) - lo + 1;
  return lo + static_cast<uint32_t>((h * span) >> 32);
}

__global__ void pair_weights_kernel(const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                                    const uint32_t* __restrict__ src, uint32_t* __restrict__ w, int64_t m,
                                    uint64_t seed, uint32_t lo, uint32_t hi) {
  int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; e < m; e += stride) w[e] = pair_weight(seed, src[e], static_cast<uint32_t>(col[e]), lo, hi);
}

// Lognormal(0,1) weight per directed edge: Box-Muller on two 32-bit draws.
__global__ void lognormal_weights_kernel(const int32_t* __restrict__ col, const uint32_t* __restrict__ src,
                                         float* __restrict__ w, int64_t m, uint64_t seed) {
  int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; e < m; e += stride) {
    uint64_t h = hash2(seed ^ 0x4C4F474E4F524D31ull, src[e], static_cast<uint32_t>(col[e]));
    double u1 = (static_cast<double>(h & 0xffffffffull) + 0.5) * (1.0 / 4294967296.0);
    double u2 = (static_cast<double>(h >> 32) + 0.5) * (1.0 / 4294967296.0);
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    w[e] = static_cast<float>(exp(z));
  }
}

// Flags any malformed CSR entry: row not monotone / out of [0,m], column id
// outside [0,n).
__global__ void validate_csr_kernel(const int32_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n,
                                    int64_t m, unsigned* __restrict__ bad) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = i; v < n; v += stride)
    if (row[v] > row[v + 1] || row[v] < 0 || row[v + 1] > m) atomicOr(bad, 1u);
  for (int64_t e = i; e < m; e += stride)
    if (col[e] < 0 || col[e] >= n) atomicOr(bad, 2u);
  if (i == 0 && (row[0] != 0 || row[n] != m)) atomicOr(bad, 4u);
}

__global__ void offset_copy_kernel(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t count,
                                   int32_t add) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < count; i += stride) out[i] = in[i] + add;
}

// ------------------------------------------------------------- host helpers

void ensure_outdeg(gx_graph* g) {
  if (g->has_outdeg) return;
  g->outdeg_buf.ensure(std::max<int64_t>(g->n, 1) * sizeof(uint32_t));
  if (g->n > 0) {
    outdeg_kernel<<<grid_for(g->n, 256, 1 << 16), 256, 0, g->stream>>>(g->row, g->outdeg_buf.as<uint32_t>(), g->n);
    GX_LAUNCHED("outdeg_kernel");
  }
  g->has_outdeg = true;
}

// Source vertex of every CSR edge (uint32[m]) into `src`.
void edge_sources(const int32_t* row, int64_t n, int64_t m, uint32_t* src, hipStream_t s) {
  if (m == 0) return;
  fill<uint32_t>(src, 0u, m, s);
  row_heads_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(row, src, n);
  GX_LAUNCHED("row_heads_kernel");
  inclusive_max_scan_u32(src, src, m, s);
}

void build_transpose(gx_graph* g) {
  if (g->has_transpose) return;
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  g->trow_buf.ensure((n + 1) * sizeof(int32_t));
  g->tcol_buf.ensure((m + kEdgePad) * sizeof(int32_t));
  g->tperm_buf.ensure(std::max<int64_t>(m, 1) * sizeof(int32_t));
  fill<int32_t>(g->tcol() + m, 0, kEdgePad, s);
  if (m == 0) {
    fill<int32_t>(g->trow(), 0, n + 1, s);
    GX_HIP(hipStreamSynchronize(s));
    g->has_transpose = true;
    return;
  }
  const int bits = bits_for(n);
  DevBuf src, k0, k1, i0, i1;
  src.ensure(m * sizeof(uint32_t));
  k0.ensure(m * sizeof(uint64_t));
  k1.ensure(m * sizeof(uint64_t));
  i0.ensure(m * sizeof(uint32_t));
  edge_sources(g->row, n, m, src.as<uint32_t>(), s);
  transpose_keys_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(g->col, src.as<uint32_t>(), k0.as<uint64_t>(),
                                                                    i0.as<uint32_t>(), m, bits);
  GX_LAUNCHED("transpose_keys_kernel");
  src.release();
  uint32_t* ids_out = reinterpret_cast<uint32_t*>(g->tperm());
  radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), i0.as<uint32_t>(), ids_out, m, 2 * bits, s);
  k0.release();
  i0.release();
  offsets_from_sorted(k1.as<uint64_t>(), m, bits, n, g->trow(), s);
  lo_bits_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(k1.as<uint64_t>(), g->tcol(), m,
                                                            (uint64_t(1) << bits) - 1);
  GX_LAUNCHED("lo_bits_kernel");
  GX_HIP(hipStreamSynchronize(s));
  g->has_transpose = true;
}

__global__ void degree_keys_kernel(const int32_t* __restrict__ row, int64_t n, uint64_t* __restrict__ keys) {
  int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; v < n; v += stride) {
    const uint32_t d = static_cast<uint32_t>(row[v + 1] - row[v]);
    keys[v] = (static_cast<uint64_t>(0x7fffffffu - d) << 31) | static_cast<uint64_t>(v);
  }
}

__global__ void rank_from_sorted_kernel(const uint64_t* __restrict__ keys, int64_t n, const int32_t* __restrict__ row,
                                        uint32_t* __restrict__ old_of_new, uint32_t* __restrict__ new_of_old,
                                        uint32_t* __restrict__ outdeg_new) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) {
    const uint32_t v = static_cast<uint32_t>(keys[i] & 0x7fffffffull);
    old_of_new[i] = v;
    new_of_old[v] = static_cast<uint32_t>(i);
    outdeg_new[i] = static_cast<uint32_t>(row[v + 1] - row[v]);
  }
}

// Relabelled in-edge keys: (new dst << bits) | new src.
__global__ void relabel_keys_kernel(const int32_t* __restrict__ col, const uint32_t* __restrict__ src,
                                    const uint32_t* __restrict__ new_of_old, uint64_t* __restrict__ keys, int64_t m,
                                    int bits) {
  int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; e < m; e += stride)
    keys[e] = (static_cast<uint64_t>(new_of_old[col[e]]) << bits) | new_of_old[src[e]];
}

// Sort `cnt` edge keys (u << bits | v), drop duplicates and self-loops, and
// make a library-owned CSR of n vertices on g.
static void csr_from_keys(gx_graph* g, DevBuf& keys, int64_t cnt, int bits, int64_t n, hipStream_t s) {
  GX_REQUIRE(cnt <= (int64_t(1) << 32), "too many generated edges for 32-bit compaction");
  DevBuf sorted;
  sorted.ensure(std::max<int64_t>(cnt, 1) * sizeof(uint64_t));
  if (cnt > 0) radix_sort_u64(keys.as<uint64_t>(), sorted.as<uint64_t>(), nullptr, nullptr, cnt, 2 * bits, s);
  keys.release();
  int64_t m = 0;
  DevBuf uniq;
  if (cnt > 0) {
    DevBuf keep, pos;
    keep.ensure(cnt * sizeof(uint32_t));
    pos.ensure(cnt * sizeof(uint32_t));
    unique_flags_kernel<<<grid_for(cnt, 256, 1 << 16), 256, 0, s>>>(sorted.as<uint64_t>(), cnt, bits,
                                                                      keep.as<uint32_t>());
    GX_LAUNCHED("unique_flags_kernel");
    exclusive_scan_u32(keep.as<uint32_t>(), pos.as<uint32_t>(), cnt, s);
    uint32_t last_pos = 0, last_keep = 0;
    GX_HIP(hipMemcpyAsync(&last_pos, pos.as<uint32_t>() + cnt - 1, 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(&last_keep, keep.as<uint32_t>() + cnt - 1, 4, hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    m = static_cast<int64_t>(last_pos) + last_keep;
    GX_REQUIRE(m < (int64_t(1) << 31), "edge count exceeds int32 CSR limit");
    uniq.ensure(std::max<int64_t>(m, 1) * sizeof(uint64_t));
    compact_kernel<<<grid_for(cnt, 256, 1 << 16), 256, 0, s>>>(sorted.as<uint64_t>(), keep.as<uint32_t>(),
                                                                 pos.as<uint32_t>(), cnt, uniq.as<uint64_t>());
    GX_LAUNCHED("compact_kernel");
  }
  sorted.release();
  g->n = n;
  g->m = m;
  g->owns_csr = true;
  g->row_buf.ensure((n + 1) * sizeof(int32_t));
  g->col_buf.ensure(std::max<int64_t>(m, 1) * sizeof(int32_t));
  g->row = g->row_buf.as<int32_t>();
  g->col = g->col_buf.as<int32_t>();
  if (m > 0) {
    offsets_from_sorted(uniq.as<uint64_t>(), m, bits, n, g->row, s);
    lo_bits_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(uniq.as<uint64_t>(), g->col, m,
                                                              (uint64_t(1) << bits) - 1);
    GX_LAUNCHED("lo_bits_kernel");
  } else {
    fill<int32_t>(g->row, 0, n + 1, s);
  }
  GX_HIP(hipStreamSynchronize(s));
}

// Relabelled out-edge keys (new src << bits | new dst) with the edge id.
__global__ void relabel_out_keys_kernel(const int32_t* __restrict__ col, const uint32_t* __restrict__ src,
                                        const uint32_t* __restrict__ new_of_old, uint64_t* __restrict__ keys,
                                        uint32_t* __restrict__ ids, int64_t m, int bits) {
  int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; e < m; e += stride) {
    keys[e] = (static_cast<uint64_t>(new_of_old[src[e]]) << bits) | new_of_old[col[e]];
    ids[e] = static_cast<uint32_t>(e);
  }
}

__global__ void gather_u32_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ idx,
                                  uint32_t* __restrict__ out, int64_t m) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < m; i += stride) out[i] = in[idx[i]];
}

void build_relabeled(gx_graph* g) {
  Relabeled& R = g->relab;
  if (R.ready) return;
  const bool weighted = (g->flags & GX_GRAPH_WEIGHTED) != 0;  // uint32 weights are carried along
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  const int bits = bits_for(n);
  R.new_of_old.ensure(std::max<int64_t>(n, 1) * 4);
  R.old_of_new.ensure(std::max<int64_t>(n, 1) * 4);
  R.row.ensure((n + 1) * 4);
  R.col.ensure(std::max<int64_t>(m, 1) * 4);
  if (weighted) R.w.ensure(std::max<int64_t>(m, 1) * 4);
  {
    DevBuf k0, k1, deg;
    k0.ensure(std::max<int64_t>(n, 1) * 8);
    k1.ensure(std::max<int64_t>(n, 1) * 8);
    deg.ensure(std::max<int64_t>(n, 1) * 4);
    degree_keys_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(g->row, n, k0.as<uint64_t>());
    GX_LAUNCHED("degree_keys_kernel");
    radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nullptr, nullptr, n, 62, s);
    rank_from_sorted_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(k1.as<uint64_t>(), n, g->row,
                                                                       R.old_of_new.as<uint32_t>(),
                                                                       R.new_of_old.as<uint32_t>(), deg.as<uint32_t>());
    GX_LAUNCHED("rank_from_sorted_kernel");
    GX_HIP(hipStreamSynchronize(s));
  }
  if (m == 0) {
    fill<int32_t>(R.row.as<int32_t>(), 0, n + 1, s);
  } else {
    DevBuf src, k0, k1, i0, perm;
    src.ensure(m * 4);
    k0.ensure(m * 8);
    i0.ensure(m * 4);
    edge_sources(g->row, n, m, src.as<uint32_t>(), s);
    relabel_out_keys_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(
        g->col, src.as<uint32_t>(), R.new_of_old.as<uint32_t>(), k0.as<uint64_t>(), i0.as<uint32_t>(), m, bits);
    GX_LAUNCHED("relabel_out_keys_kernel");
    GX_HIP(hipStreamSynchronize(s));
    src.release();
    k1.ensure(m * 8);
    perm.ensure(m * 4);
    radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), i0.as<uint32_t>(), perm.as<uint32_t>(), m, 2 * bits, s);
    k0.release();
    i0.release();
    offsets_from_sorted(k1.as<uint64_t>(), m, bits, n,
                                                                          R.row.as<int32_t>(), s);
    lo_bits_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(k1.as<uint64_t>(), R.col.as<int32_t>(), m,
                                                              (uint64_t(1) << bits) - 1);
    GX_LAUNCHED("lo_bits_kernel");
    if (weighted) {
      gather_u32_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(static_cast<const uint32_t*>(g->w),
                                                                   perm.as<uint32_t>(), R.w.as<uint32_t>(), m);
      GX_LAUNCHED("gather_u32_kernel");
    }
  }
  GX_HIP(hipStreamSynchronize(s));
  R.ready = true;
}

__global__ void unpermute_u32_kernel(const uint32_t* __restrict__ in, const uint32_t* __restrict__ new_of_old,
                                     uint32_t* __restrict__ out, int64_t n) {
  for (int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[v] = in[new_of_old[v]];
}

void unpermute_u32(const uint32_t* in, const uint32_t* new_of_old, uint32_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  unpermute_u32_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(in, new_of_old, out, n);
  GX_LAUNCHED("unpermute_u32_kernel");
}

void build_degree_order(gx_graph* g) {
  DegreeOrder& D = g->degree_order;
  if (D.ready) return;
  const int64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  const int bits = bits_for(n);
  D.new_of_old.ensure(std::max<int64_t>(n, 1) * 4);
  D.old_of_new.ensure(std::max<int64_t>(n, 1) * 4);
  D.outdeg.ensure(std::max<int64_t>(n, 1) * 4);
  D.trow.ensure((n + 1) * 4);
  D.tcol.ensure((m + kEdgePad) * 4);
  fill<int32_t>(D.tcol.as<int32_t>() + m, 0, kEdgePad, s);
  {
    DevBuf k0, k1;
    k0.ensure(std::max<int64_t>(n, 1) * 8);
    k1.ensure(std::max<int64_t>(n, 1) * 8);
    degree_keys_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(g->row, n, k0.as<uint64_t>());
    GX_LAUNCHED("degree_keys_kernel");
    radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nullptr, nullptr, n, 62, s);
    rank_from_sorted_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(k1.as<uint64_t>(), n, g->row,
                                                                       D.old_of_new.as<uint32_t>(),
                                                                       D.new_of_old.as<uint32_t>(),
                                                                       D.outdeg.as<uint32_t>());
    GX_LAUNCHED("rank_from_sorted_kernel");
  }
  if (m == 0) {
    fill<int32_t>(D.trow.as<int32_t>(), 0, n + 1, s);
  } else {
    DevBuf src, k0, k1;
    src.ensure(m * 4);
    k0.ensure(m * 8);
    k1.ensure(m * 8);
    edge_sources(g->row, n, m, src.as<uint32_t>(), s);
    relabel_keys_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(g->col, src.as<uint32_t>(),
                                                                   D.new_of_old.as<uint32_t>(), k0.as<uint64_t>(), m,
                                                                   bits);
    GX_LAUNCHED("relabel_keys_kernel");
    src.release();
    radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), nullptr, nullptr, m, 2 * bits, s);
    k0.release();
    offsets_from_sorted(k1.as<uint64_t>(), m, bits, n,
                                                                          D.trow.as<int32_t>(), s);
    lo_bits_kernel<<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(k1.as<uint64_t>(), D.tcol.as<int32_t>(), m,
                                                              (uint64_t(1) << bits) - 1);
    GX_LAUNCHED("lo_bits_kernel");
  }
  GX_HIP(hipStreamSynchronize(s));
  D.ready = true;
}

static void add_pair_weights(gx_graph* g, uint64_t seed, uint32_t lo, uint32_t hi) {
  hipStream_t s = g->stream;
  g->w_buf.ensure(std::max<int64_t>(g->m, 1) * sizeof(uint32_t));
  g->w = g->w_buf.p;
  g->flags |= GX_GRAPH_WEIGHTED;
  if (g->m == 0) return;
  DevBuf src;
  src.ensure(g->m * sizeof(uint32_t));
  edge_sources(g->row, g->n, g->m, src.as<uint32_t>(), s);
  pair_weights_kernel<<<grid_for(g->m, 256, 1 << 16), 256, 0, s>>>(g->row, g->col, src.as<uint32_t>(),
                                                                     g->w_buf.as<uint32_t>(), g->m, seed, lo, hi);
  GX_LAUNCHED("pair_weights_kernel");
  GX_HIP(hipStreamSynchronize(s));
}

static void add_lognormal_weights(gx_graph* g, uint64_t seed) {
  hipStream_t s = g->stream;
  g->w_buf.ensure(std::max<int64_t>(g->m, 1) * sizeof(float));
  g->w = g->w_buf.p;
  g->flags |= GX_GRAPH_FWEIGHTED;
  if (g->m == 0) return;
  DevBuf src;
  src.ensure(g->m * sizeof(uint32_t));
  edge_sources(g->row, g->n, g->m, src.as<uint32_t>(), s);
  lognormal_weights_kernel<<<grid_for(g->m, 256, 1 << 16), 256, 0, s>>>(g->col, src.as<uint32_t>(),
                                                                          g->w_buf.as<float>(), g->m, seed);
  GX_LAUNCHED("lognormal_weights_kernel");
  GX_HIP(hipStreamSynchronize(s));
}

}  // namespace gx

// ------------------------------------------------------------------ C ABI

extern "C" gx_status gx_graph_create(int64_t n, int64_t m, const int32_t* row, const int32_t* col,
                                     const void* weights, uint32_t flags, gx_graph** out) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  GX_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "vertex count out of range");
  GX_REQUIRE(m >= 0 && m < (int64_t(1) << 31), "edge count out of range");
  check_ptr(row, "row");
  GX_REQUIRE(m == 0 || col != nullptr, "col is NULL");
  GX_REQUIRE(row[0] == 0 && row[n] == m, "row offsets must start at 0 and end at m");
  for (int64_t v = 0; v < n; ++v) GX_REQUIRE(row[v] <= row[v + 1], "row offsets must be non-decreasing");
  for (int64_t e = 0; e < m; ++e) GX_REQUIRE(col[e] >= 0 && col[e] < n, "column index out of range");
  if (flags & (GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED)) GX_REQUIRE(m == 0 || weights, "weights flagged but NULL");
  auto g = std::make_unique<gx_graph>();
  g->n = n;
  g->m = m;
  g->flags = flags;
  g->row_buf.ensure((n + 1) * sizeof(int32_t));
  g->col_buf.ensure(std::max<int64_t>(m, 1) * sizeof(int32_t));
  g->row = g->row_buf.as<int32_t>();
  g->col = g->col_buf.as<int32_t>();
  GX_HIP(hipMemcpy(g->row, row, (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  if (m) GX_HIP(hipMemcpy(g->col, col, m * sizeof(int32_t), hipMemcpyHostToDevice));
  if (flags & (GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED)) {
    g->w_buf.ensure(std::max<int64_t>(m, 1) * 4);
    g->w = g->w_buf.p;
    if (m) GX_HIP(hipMemcpy(g->w, weights, m * 4, hipMemcpyHostToDevice));
  }
  *out = g.release();
  GX_API_END
}

extern "C" gx_status gx_graph_create_device(int64_t n, int64_t m, int32_t* row_dev, int32_t* col_dev,
                                            void* weights_dev, uint32_t flags, gx_graph** out) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  GX_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "vertex count out of range");
  GX_REQUIRE(m >= 0 && m < (int64_t(1) << 31), "edge count out of range");
  check_ptr(row_dev, "row");
  GX_REQUIRE(m == 0 || col_dev != nullptr, "col is NULL");
  if (flags & (GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED)) GX_REQUIRE(m == 0 || weights_dev, "weights flagged but NULL");
  auto g = std::make_unique<gx_graph>();
  g->n = n;
  g->m = m;
  g->flags = flags;
  g->row = row_dev;
  g->col = col_dev;
  g->w = weights_dev;
  g->owns_csr = false;
  {
    DevBuf bad;
    bad.ensure(sizeof(unsigned));
    GX_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), g->stream));
    validate_csr_kernel<<<grid_for(std::max(n, m), 256, 4096), 256, 0, g->stream>>>(row_dev, col_dev, n, m,
                                                                                     bad.as<unsigned>());
    GX_LAUNCHED("validate_csr_kernel");
    unsigned hbad = 0;
    GX_HIP(hipMemcpy(&hbad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost));
    GX_REQUIRE(hbad == 0, "malformed device CSR (row offsets or column ids out of range)");
  }
  *out = g.release();
  GX_API_END
}

extern "C" gx_status gx_graph_free(gx_graph* g) {
  GX_API_BEGIN
  delete g;
  GX_API_END
}

extern "C" gx_status gx_graph_info(const gx_graph* g, int64_t* n, int64_t* m, uint32_t* flags) {
  GX_API_BEGIN
  check_graph(g);
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (flags) *flags = g->flags;
  GX_API_END
}

extern "C" gx_status gx_graph_set_stream(gx_graph* g, void* stream) {
  GX_API_BEGIN
  check_graph(g);
  g->stream = static_cast<hipStream_t>(stream);
  GX_API_END
}

extern "C" gx_status gx_graph_device_arrays(const gx_graph* g, const int32_t** row, const int32_t** col,
                                            const void** weights, const int32_t** trow, const int32_t** tcol,
                                            const int32_t** tperm) {
  GX_API_BEGIN
  check_graph(g);
  if (row) *row = g->row;
  if (col) *col = g->col;
  if (weights) *weights = g->w;
  if (trow) *trow = g->has_transpose ? g->trow() : nullptr;
  if (tcol) *tcol = g->has_transpose ? g->tcol() : nullptr;
  if (tperm) *tperm = g->has_transpose ? g->tperm() : nullptr;
  GX_API_END
}

extern "C" gx_status gx_graph_download(const gx_graph* g, int32_t* row, int32_t* col, void* weights) {
  GX_API_BEGIN
  check_graph(g);
  if (row) GX_HIP(hipMemcpy(row, g->row, (g->n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (col && g->m) GX_HIP(hipMemcpy(col, g->col, g->m * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (weights && g->m) {
    GX_REQUIRE(g->w != nullptr, "graph has no weights");
    GX_HIP(hipMemcpy(weights, g->w, g->m * 4, hipMemcpyDeviceToHost));
  }
  GX_API_END
}

extern "C" gx_status gx_graph_build_transpose(gx_graph* g) {
  GX_API_BEGIN
  check_graph(g);
  build_transpose(g);
  GX_API_END
}

extern "C" gx_status gx_graph_download_transpose(gx_graph* g, int32_t* trow, int32_t* tcol, int32_t* tperm) {
  GX_API_BEGIN
  check_graph(g);
  build_transpose(g);
  if (trow) GX_HIP(hipMemcpy(trow, g->trow(), (g->n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (tcol && g->m) GX_HIP(hipMemcpy(tcol, g->tcol(), g->m * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (tperm && g->m) GX_HIP(hipMemcpy(tperm, g->tperm(), g->m * sizeof(int32_t), hipMemcpyDeviceToHost));
  GX_API_END
}

static uint64_t prob_to_u32_threshold(double p) {
  if (p <= 0) return 0;
  if (p >= 1) return uint64_t(1) << 32;
  return static_cast<uint64_t>(std::floor(p * 4294967296.0));
}

extern "C" gx_status gx_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                                 int symmetric, int weight_kind, uint32_t wlo, uint32_t whi, gx_graph** out) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  GX_REQUIRE(scale >= 1 && scale <= 30, "RMAT scale must be in [1,30]");
  GX_REQUIRE(edge_factor >= 1, "edge factor must be >= 1");
  GX_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "RMAT probabilities invalid");
  GX_REQUIRE(weight_kind >= 0 && weight_kind <= 2, "weight_kind must be 0, 1 or 2");
  if (weight_kind == 1) GX_REQUIRE(wlo >= 1 && whi >= wlo, "weight range invalid");
  const int64_t n = int64_t(1) << scale;
  const int64_t mgen = n * edge_factor;
  const int64_t cnt = symmetric ? 2 * mgen : mgen;
  auto g = std::make_unique<gx_graph>();
  hipStream_t s = g->stream;
  DevBuf keys;
  keys.ensure(cnt * sizeof(uint64_t));
  uint64_t ta = prob_to_u32_threshold(a), tab = prob_to_u32_threshold(a + b), tabc = prob_to_u32_threshold(a + b + c);
  rmat_keys_kernel<<<grid_for(mgen, 256, 1 << 16), 256, 0, s>>>(keys.as<uint64_t>(), mgen, static_cast<int>(scale), ta,
                                                                 tab, tabc, seed, symmetric);
  GX_LAUNCHED("rmat_keys_kernel");
  csr_from_keys(g.get(), keys, cnt, static_cast<int>(scale), n, s);
  g->flags = symmetric ? GX_GRAPH_SYMMETRIC : 0u;
  if (weight_kind == 1) add_pair_weights(g.get(), seed, wlo, whi);
  if (weight_kind == 2) add_lognormal_weights(g.get(), seed);
  *out = g.release();
  GX_API_END
}

extern "C" gx_status gx_gen_grid(uint32_t side, uint32_t drop_ppm, uint32_t shortcut_ppm, uint64_t seed, uint32_t wlo,
                                 uint32_t whi, gx_graph** out) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  GX_REQUIRE(side >= 2 && side <= 32768, "grid side must be in [2, 32768]");
  GX_REQUIRE(drop_ppm <= 1000000, "drop_ppm must be <= 1e6");
  GX_REQUIRE(wlo >= 1 && whi >= wlo, "weight range invalid");
  const int64_t n = int64_t(side) * side;
  const int64_t nlat = 2 * int64_t(side) * (side - 1);
  const int64_t nshort = static_cast<int64_t>((static_cast<double>(nlat) * shortcut_ppm) / 1e6 + 0.5);
  const int64_t cnt = 2 * (nlat + nshort);
  const int bits = bits_for(n);
  auto g = std::make_unique<gx_graph>();
  hipStream_t s = g->stream;
  DevBuf keys;
  keys.ensure(cnt * sizeof(uint64_t));
  uint64_t drop_thresh = (static_cast<uint64_t>(drop_ppm) << 32) / 1000000ull;
  grid_keys_kernel<<<grid_for(nlat + nshort, 256, 1 << 16), 256, 0, s>>>(keys.as<uint64_t>(), side, nlat, nshort,
                                                                          drop_thresh, seed, bits);
  GX_LAUNCHED("grid_keys_kernel");
  csr_from_keys(g.get(), keys, cnt, bits, n, s);
  g->flags = GX_GRAPH_SYMMETRIC;
  add_pair_weights(g.get(), seed, wlo, whi);
  *out = g.release();
  GX_API_END
}

extern "C" gx_status gx_graph_union(gx_graph* const* parts, uint32_t count, gx_graph** out, int64_t* vertex_offsets) {
  GX_API_BEGIN
  check_ptr(out, "out");
  *out = nullptr;
  GX_REQUIRE(count >= 1 && parts != nullptr, "need at least one graph");
  int64_t n = 0, m = 0;
  uint32_t flags = parts[0]->flags;
  for (uint32_t i = 0; i < count; ++i) {
    check_graph(parts[i]);
    if (vertex_offsets) vertex_offsets[i] = n;
    n += parts[i]->n;
    m += parts[i]->m;
    flags &= parts[i]->flags;
  }
  if (vertex_offsets) vertex_offsets[count] = n;
  GX_REQUIRE(n < (int64_t(1) << 31) && m < (int64_t(1) << 31), "union too large for int32 CSR");
  auto g = std::make_unique<gx_graph>();
  hipStream_t s = g->stream;
  g->n = n;
  g->m = m;
  g->flags = flags & (GX_GRAPH_SYMMETRIC | GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED);
  g->row_buf.ensure((n + 1) * sizeof(int32_t));
  g->col_buf.ensure(std::max<int64_t>(m, 1) * sizeof(int32_t));
  g->row = g->row_buf.as<int32_t>();
  g->col = g->col_buf.as<int32_t>();
  bool weighted = (g->flags & (GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED)) != 0;
  if (weighted) {
    g->w_buf.ensure(std::max<int64_t>(m, 1) * 4);
    g->w = g->w_buf.p;
  }
  int64_t vo = 0, eo = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const gx_graph* p = parts[i];
    if (p->n > 0) {
      offset_copy_kernel<<<grid_for(p->n, 256, 1 << 16), 256, 0, s>>>(p->row, g->row + vo, p->n,
                                                                        static_cast<int32_t>(eo));
      GX_LAUNCHED("offset_copy_kernel");
    }
    if (p->m > 0) {
      offset_copy_kernel<<<grid_for(p->m, 256, 1 << 16), 256, 0, s>>>(p->col, g->col + eo, p->m,
                                                                        static_cast<int32_t>(vo));
      GX_LAUNCHED("offset_copy_kernel");
      if (weighted) GX_HIP(hipMemcpyAsync(static_cast<char*>(g->w) + eo * 4, p->w, p->m * 4,
                                          hipMemcpyDeviceToDevice, s));
    }
    vo += p->n;
    eo += p->m;
  }
  int32_t last = static_cast<int32_t>(m);
  GX_HIP(hipMemcpyAsync(g->row + n, &last, sizeof(int32_t), hipMemcpyHostToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  *out = g.release();
  GX_API_END
}
