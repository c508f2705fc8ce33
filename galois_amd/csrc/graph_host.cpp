// Host-side pieces of the C ABI: error reporting, device init, the .gr file
// reader/writer and the lonestar source picker.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "gx_internal.hpp"

namespace gx {

static thread_local std::string g_last_error;
static std::atomic<int> g_profiling{0};

void set_last_error(const std::string& msg) { g_last_error = msg; }
bool profiling_enabled() { return g_profiling.load() != 0; }

}  // namespace gx

using namespace gx;

extern "C" size_t gx_last_error(char* buf, size_t len) {
  const std::string& s = g_last_error;
  if (buf && len) {
    size_t k = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
  }
  return s.size();
}

extern "C" void gx_set_profiling(int on) { g_profiling.store(on ? 1 : 0); }

namespace gx {
namespace {
// Built-in defaults, which may depend on the graph's size.
int64_t default_hub_degree(int64_t, int64_t) { return 128; }
int64_t default_hot_sources(int64_t, int64_t) { return 0; }
int64_t default_blocking(int64_t, int64_t) { return 1; }
int64_t default_visited_bitmap(int64_t n, int64_t) { return n >= (int64_t(1) << 25) ? 1 : 0; }
int64_t default_max_merge(int64_t, int64_t) { return 256; }
int64_t default_bc_chunk(int64_t, int64_t) { return kEdgeChunk; }
int64_t default_off(int64_t, int64_t) { return 0; }
// The degree-ordered copy pays from 2^24 vertices on (RMAT-26 SSSP 72.6 ->
// 49.5 ms, BFS 17.2 -> 15.5 ms); on RMAT-18 BFS it cost 0.41 -> 0.44 ms and
// on RMAT-22 SSSP nothing changed.
int64_t default_relabel(int64_t n, int64_t) { return n >= (int64_t(1) << 24) ? 1 : 0; }
// Hubs expanded in distance order, one distance at a time (r = 1), in the
// relax steps whenever no near item is closer: RMAT-26 SSSP 49.7 -> 39.0 ms
// (1.84 -> 1.10 relaxations per edge; two distances 39.8 ms); on RMAT-22 the
// extra steps cost more than the relaxations saved (5.7 -> 6.7 ms).
int64_t default_hub_refine(int64_t n, int64_t) { return n >= (int64_t(1) << 24) ? 1 : 0; }
int64_t default_batch_width(int64_t, int64_t) { return 64; }
int64_t default_bc_grid(int64_t, int64_t) { return 256 * 32; }
struct Knob {
  const char* key;
  std::atomic<int64_t> value;
  int64_t lo, hi;
  int64_t (*dflt)(int64_t n, int64_t m);
};
Knob g_knobs[] = {
    {"frontier.hub_degree", {-1}, 64, int64_t(1) << 30, default_hub_degree},
    {"pr.hot_sources", {-1}, 0, int64_t(1) << 31, default_hot_sources},
    {"pr.blocking", {-1}, 0, 2, default_blocking},
    {"pr.task_edges", {-1}, int64_t(1) << 12, int64_t(1) << 30, pr_task_edges_default},
    {"bfs.visited_bitmap", {-1}, 0, 1, default_visited_bitmap},
    {"sssp.max_merge", {-1}, 1, int64_t(1) << 16, default_max_merge},
    {"sssp.defer_hubs", {-1}, 0, 1, default_off},
    {"sssp.hub_refine", {-1}, 0, 32, default_hub_refine},
    {"sssp.batch_width", {-1}, 0, int64_t(1) << 24, default_batch_width},
    {"frontier.relabel", {-1}, 0, 1, default_relabel},
    {"bc.chunk", {-1}, 64, 4096, default_bc_chunk},
    {"bc.grid", {-1}, 256, 65536, default_bc_grid},
};
Knob* find_knob(const char* key) {
  if (!key) return nullptr;
  for (Knob& k : g_knobs)
    if (std::strcmp(k.key, key) == 0) return &k;
  return nullptr;
}
}  // namespace

int64_t pr_task_edges_default(int64_t, int64_t m) {
  constexpr int64_t lo = int64_t(1) << 19, hi = int64_t(1) << 20;
  return std::min(std::max(m >> 10, lo), hi);
}

int64_t tune_effective(const char* key, int64_t n, int64_t m) {
  Knob* k = find_knob(key);
  if (!k) return -1;
  const int64_t v = k->value.load();
  return v < 0 ? k->dflt(n, m) : v;
}
}  // namespace gx

extern "C" gx_status gx_tune_set(const char* key, int64_t value) {
  GX_API_BEGIN
  Knob* k = find_knob(key);
  GX_REQUIRE(k, std::string("unknown tuning key ") + (key ? key : "(null)"));
  GX_REQUIRE(value == -1 || (value >= k->lo && value <= k->hi), std::string("value out of range for ") + key);
  k->value.store(value);
  GX_API_END
}

extern "C" gx_status gx_tune_get(const char* key, int64_t* value) {
  GX_API_BEGIN
  Knob* k = find_knob(key);
  GX_REQUIRE(k, std::string("unknown tuning key ") + (key ? key : "(null)"));
  check_ptr(value, "value");
  *value = k->value.load();
  GX_API_END
}

extern "C" gx_status gx_tune_effective(const char* key, int64_t n, int64_t m, int64_t* value) {
  GX_API_BEGIN
  Knob* k = find_knob(key);
  GX_REQUIRE(k, std::string("unknown tuning key ") + (key ? key : "(null)"));
  check_ptr(value, "value");
  GX_REQUIRE(n >= 0 && m >= 0, "negative graph size");
  *value = tune_effective(key, n, m);
  GX_API_END
}

extern "C" gx_status gx_init(int device) {
  GX_API_BEGIN
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) fail(GX_ERR_NO_DEVICE, "no HIP device visible");
  GX_REQUIRE(device >= 0 && device < count, "device index out of range");
  GX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GX_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    fail(GX_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library is built for gfx950");
  (void)noalloc_scratch();  // this device's copy of gx_cc_noalloc's scratch
  GX_API_END
}

// ------------------------------------------------------------------ .gr I/O
// Layout (libgalois/src/FileGraph.cpp:44-53): four little-endian uint64
// {version, sizeof edge datum, n, m}; n uint64 END offsets (row[v+1]); m
// destinations (uint32 for version 1, uint64 for version 2); for version 1 a
// 4-byte pad when m is odd; then m edge data of `sizeof edge datum` bytes.
// The reader skips `m + m % 2` destination slots for both versions, as
// FileGraph::fromMem does (FileGraph.cpp:121-139).

namespace {
struct FileCloser {
  void operator()(FILE* f) const { if (f) std::fclose(f); }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

void read_exact(FILE* f, void* dst, size_t bytes, const char* what) {
  if (bytes && std::fread(dst, 1, bytes, f) != bytes)
    fail(GX_ERR_IO, std::string("truncated .gr file while reading ") + what);
}
void write_exact(FILE* f, const void* src, size_t bytes) {
  if (bytes && std::fwrite(src, 1, bytes, f) != bytes) fail(GX_ERR_IO, "short write to .gr file");
}
}  // namespace

extern "C" void gx_host_csr_free(gx_host_csr* g) {
  if (!g) return;
  std::free(g->row);
  std::free(g->col);
  std::free(g->edge_data);
  g->row = nullptr;
  g->col = nullptr;
  g->edge_data = nullptr;
  g->n = g->m = 0;
}

extern "C" gx_status gx_gr_read(const char* path, gx_host_csr* out) {
  GX_API_BEGIN
  check_ptr(path, "path");
  check_ptr(out, "out");
  std::memset(out, 0, sizeof(*out));
  FilePtr f(std::fopen(path, "rb"));
  if (!f) fail(GX_ERR_IO, std::string("cannot open ") + path);
  uint64_t hdr[4];
  read_exact(f.get(), hdr, sizeof(hdr), "header");
  const uint64_t ver = hdr[0], esz = hdr[1], n = hdr[2], m = hdr[3];
  if (ver != 1 && ver != 2) fail(GX_ERR_IO, "unknown .gr version " + std::to_string(ver));
  if (n >= (uint64_t(1) << 31) || m >= (uint64_t(1) << 31))
    fail(GX_ERR_IO, "graph too large for the int32 CSR layout");
  if (esz > 8) fail(GX_ERR_IO, "edge datum size " + std::to_string(esz) + " unsupported");
  {
    // The header sizes every allocation below: a file shorter than the
    // sections its header announces is rejected before anything is allocated.
    const uint64_t dsz = ver == 1 ? 4 : 8;
    const uint64_t need = sizeof(hdr) + 8 * n + dsz * (m + (m % 2)) + esz * m;
    if (std::fseek(f.get(), 0, SEEK_END) != 0) fail(GX_ERR_IO, std::string("cannot seek in ") + path);
    const long have = std::ftell(f.get());
    if (have < 0 || static_cast<uint64_t>(have) < need)
      fail(GX_ERR_IO, "file is shorter than its header announces (" + std::to_string(have < 0 ? 0 : have) +
                          " of " + std::to_string(need) + " bytes)");
    if (std::fseek(f.get(), sizeof(hdr), SEEK_SET) != 0) fail(GX_ERR_IO, std::string("cannot seek in ") + path);
  }
  gx_host_csr g{};
  g.n = static_cast<int64_t>(n);
  g.m = static_cast<int64_t>(m);
  g.edge_size = static_cast<uint32_t>(esz);
  g.version = static_cast<uint32_t>(ver);
  g.row = static_cast<int32_t*>(std::malloc((n + 1) * sizeof(int32_t)));
  g.col = static_cast<int32_t*>(std::malloc(std::max<uint64_t>(m, 1) * sizeof(int32_t)));
  if (!g.row || !g.col) {
    gx_host_csr_free(&g);
    fail(GX_ERR_ALLOC, "host allocation failed");
  }
  try {
    std::vector<uint64_t> idx(n);
    read_exact(f.get(), idx.data(), n * 8, "offsets");
    g.row[0] = 0;
    uint64_t prev = 0;
    for (uint64_t v = 0; v < n; ++v) {
      if (idx[v] < prev || idx[v] > m) fail(GX_ERR_IO, "non-monotone or out-of-range edge offsets");
      g.row[v + 1] = static_cast<int32_t>(idx[v]);
      prev = idx[v];
    }
    if (n > 0 && idx[n - 1] != m) fail(GX_ERR_IO, "last edge offset does not equal the edge count");
    if (n == 0 && m != 0) fail(GX_ERR_IO, "edges without vertices");
    if (ver == 1) {
      read_exact(f.get(), g.col, m * 4, "destinations");
      if (m % 2) {
        uint32_t pad;
        read_exact(f.get(), &pad, 4, "padding");
      }
    } else {
      std::vector<uint64_t> d(m);
      read_exact(f.get(), d.data(), m * 8, "destinations");
      if (m % 2) {
        uint64_t pad;
        read_exact(f.get(), &pad, 8, "padding");
      }
      for (uint64_t e = 0; e < m; ++e) g.col[e] = static_cast<int32_t>(d[e]);
    }
    for (uint64_t e = 0; e < m; ++e)
      if (static_cast<uint32_t>(g.col[e]) >= n) fail(GX_ERR_IO, "edge destination out of range");
    if (esz && m) {
      g.edge_data = std::malloc(m * esz);
      if (!g.edge_data) fail(GX_ERR_ALLOC, "host allocation failed");
      read_exact(f.get(), g.edge_data, m * esz, "edge data");
    }
  } catch (...) {
    gx_host_csr_free(&g);
    throw;
  }
  *out = g;
  GX_API_END
}

extern "C" gx_status gx_gr_write(const char* path, const gx_host_csr* g) {
  GX_API_BEGIN
  check_ptr(path, "path");
  check_ptr(g, "graph");
  GX_REQUIRE(g->n >= 0 && g->m >= 0, "negative sizes");
  check_ptr(g->row, "row");
  GX_REQUIRE(g->m == 0 || g->col, "col is NULL");
  GX_REQUIRE(g->edge_size == 0 || g->m == 0 || g->edge_data, "edge data is NULL");
  GX_REQUIRE(g->row[0] == 0 && g->row[g->n] == g->m, "row offsets must start at 0 and end at m");
  FilePtr f(std::fopen(path, "wb"));
  if (!f) fail(GX_ERR_IO, std::string("cannot create ") + path);
  const uint64_t n = g->n, m = g->m;
  uint64_t hdr[4] = {1, g->edge_size, n, m};
  write_exact(f.get(), hdr, sizeof(hdr));
  std::vector<uint64_t> idx(n);
  for (uint64_t v = 0; v < n; ++v) idx[v] = static_cast<uint64_t>(g->row[v + 1]);
  write_exact(f.get(), idx.data(), n * 8);
  write_exact(f.get(), g->col, m * 4);
  if (m % 2) {
    uint32_t pad = 0;
    write_exact(f.get(), &pad, 4);
  }
  if (g->edge_size && m) write_exact(f.get(), g->edge_data, m * g->edge_size);
  if (std::fflush(f.get()) != 0) fail(GX_ERR_IO, "flush failed");
  GX_API_END
}

// ------------------------------------------------------------ source picker
// lonestar/liblonestar/include/Lonestar/Utils.h:12-31: std::mt19937 seeded
// with kRandSeed, std::uniform_int_distribution over [0, n-1], rejecting
// vertices of degree 0.  The same standard-library engine and distribution
// are used so the sequence matches the reference built with the same library.
extern "C" gx_status gx_pick_sources(int64_t n, const int32_t* row, uint32_t count, uint32_t seed, uint32_t* out) {
  GX_API_BEGIN
  check_ptr(row, "row");
  GX_REQUIRE(count == 0 || out != nullptr, "out is NULL");
  GX_REQUIRE(n > 0, "graph has no vertices");
  bool any = false;
  for (int64_t v = 0; v < n && !any; ++v) any = row[v + 1] > row[v];
  GX_REQUIRE(any || count == 0, "graph has no vertex with edges");
  std::mt19937 rng(seed);
  std::uniform_int_distribution<uint32_t> udist(0, static_cast<uint32_t>(n - 1));
  for (uint32_t i = 0; i < count; ++i) {
    uint32_t s;
    do {
      s = udist(rng);
    } while (row[s + 1] == row[s]);
    out[i] = s;
  }
  GX_API_END
}

// isApproximateDegreeDistributionPowerLaw (lonestar/liblonestar/include/
// Lonestar/Utils.h:37-57): average degree m/n (integer) below 10 is not a
// power law; otherwise the degrees of 1000 vertices (n if fewer) drawn by
// the source picker (seed 27491095, degree-0 vertices rejected) are sorted,
// and the graph is a power law when their mean / 1.25 exceeds their median
// (the element at samples/2).  The sum is a uint32 as in the reference.
extern "C" gx_status gx_degree_power_law(int64_t n, int64_t m, const int32_t* row, int* power_law) {
  GX_API_BEGIN
  check_ptr(row, "row");
  check_ptr(power_law, "power_law");
  GX_REQUIRE(n > 0, "graph has no vertices");
  GX_REQUIRE(m >= 0 && row[n] == m, "row[n] differs from m");
  *power_law = 0;
  const uint32_t average = static_cast<uint32_t>(m / n);
  if (average < 10) return GX_OK;
  const uint32_t samples = n < 1000 ? static_cast<uint32_t>(n) : 1000u;
  std::vector<uint32_t> pick(samples), deg(samples);
  const gx_status st = gx_pick_sources(n, row, samples, 27491095u, pick.data());
  if (st != GX_OK) return st;
  uint32_t total = 0;
  for (uint32_t i = 0; i < samples; ++i) {
    deg[i] = static_cast<uint32_t>(row[pick[i] + 1] - row[pick[i]]);
    total += deg[i];
  }
  std::sort(deg.begin(), deg.end());
  const double mean = static_cast<double>(total) / samples;
  const double median = deg[samples / 2];
  *power_law = mean / 1.25 > median ? 1 : 0;
  GX_API_END
}
