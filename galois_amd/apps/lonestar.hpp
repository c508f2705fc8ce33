// Lonestar-compatible front end for the app entry points: the command line
// (llvm::cl syntax: -name=value, -name value, --name, bare boolean -flag,
// one positional input file), the banner and the statistics table of
// lonestar/liblonestar/src/BoilerPlate.cpp:25-96 (-t, -noverify, -statFile,
// -symmetricGraph), libgalois/src/Statistics.cpp:78-92 and
// libgalois/include/galois/runtime/Statistics.h:340-440 (header
// "STAT_TYPE, REGION, CATEGORY, TOTAL_TYPE, TOTAL"; StatTimer values in
// milliseconds as TMAX), and the INFO/ERROR line prefixes of
// libgalois/src/gIO.cpp:37-47.  The apps call the GPU kernels through the
// C ABI of include/galois_amd.h only.
#pragma once

#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "galois_amd.h"

namespace lonestar {

[[noreturn]] inline void die(const std::string& msg) {
  std::cerr << "ERROR: " << msg << "\n";
  std::exit(1);
}

inline void check(gx_status s, const char* what) {
  if (s == GX_OK) return;
  char buf[4096];
  gx_last_error(buf, sizeof(buf));
  die(std::string(what) + ": " + buf);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

template <class... A>
void info(const A&... a) {
  std::ostringstream os;
  (os << ... << a);
  std::cout << "INFO: " << os.str() << "\n";
}

template <class... A>
void print(const A&... a) {
  (std::cout << ... << a);
}

// ------------------------------------------------------------ command line
class Cli {
 public:
  Cli() {
    flag("noverify", "Skip verification step (default value false)");
    value("t", "1", "Number of threads (default value 1)");
    value("statFile", "", "ouput file to print stats to (default value empty)");
    flag("symmetricGraph", "Specify that the input graph is symmetric");
    value("o", "", "Write the per-node result to this file, one 'node value' line per node (- = stdout)");
  }
  void flag(const std::string& name, const std::string& desc) { opts_[name] = Opt{"false", desc, true, false}; }
  void value(const std::string& name, const std::string& def, const std::string& desc) {
    opts_[name] = Opt{def, desc, false, false};
  }
  // Parses argv; the single positional argument is the input file.
  void parse(int argc, char** argv, const char* app) {
    app_ = app;
    for (int i = 0; i < argc; ++i) cmdline_ += (i ? " " : "") + std::string(argv[i]);
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      if (a.size() > 1 && a[0] == '-') {
        std::string body = a.substr(a[1] == '-' ? 2 : 1);
        if (body == "help" || body == "h") usage();
        std::string name = body, val;
        bool has_val = false;
        const size_t eq = body.find('=');
        if (eq != std::string::npos) {
          name = body.substr(0, eq);
          val = body.substr(eq + 1);
          has_val = true;
        }
        auto it = opts_.find(name);
        if (it == opts_.end()) die(std::string(app) + ": Unknown command line argument '" + a + "'");
        Opt& o = it->second;
        if (o.is_bool) {
          o.val = has_val ? val : "true";
        } else if (has_val) {
          o.val = val;
        } else {
          if (i + 1 >= argc) die(std::string(app) + ": option '" + name + "' requires a value");
          o.val = argv[++i];
        }
        o.set = true;
      } else {
        if (!input_.empty()) die(std::string(app) + ": too many positional arguments");
        input_ = a;
      }
    }
    if (input_.empty()) die(std::string(app) + ": Not enough positional command line arguments specified!");
  }
  std::string str(const std::string& n) const { return get(n).val; }
  bool on(const std::string& n) const {
    const std::string& v = get(n).val;
    return v == "true" || v == "1" || v == "TRUE" || v == "True";
  }
  uint64_t u64(const std::string& n) const { return std::strtoull(get(n).val.c_str(), nullptr, 0); }
  double f64(const std::string& n) const { return std::strtod(get(n).val.c_str(), nullptr); }
  bool given(const std::string& n) const { return get(n).set; }
  const std::string& input() const { return input_; }
  const std::string& cmdline() const { return cmdline_; }

 private:
  struct Opt {
    std::string val, desc;
    bool is_bool, set;
  };
  const Opt& get(const std::string& n) const {
    auto it = opts_.find(n);
    if (it == opts_.end()) die("internal: option " + n + " not declared");
    return it->second;
  }
  [[noreturn]] void usage() const {
    std::cout << "USAGE: " << app_ << " [options] <input file>\n\nOPTIONS:\n";
    for (const auto& kv : opts_) std::cout << "  -" << kv.first << " - " << kv.second.desc << "\n";
    std::exit(0);
  }
  std::map<std::string, Opt> opts_;
  std::string input_, cmdline_, app_;
};

// -------------------------------------------------------------- statistics
class Stats {
 public:
  void stat_int(const std::string& region, const std::string& cat, const char* total, int64_t v) {
    ints_.push_back({region, cat, total, std::to_string(v)});
  }
  void stat_fp(const std::string& region, const std::string& cat, const char* total, double v) {
    std::ostringstream os;
    os << v;
    fps_.push_back({region, cat, total, os.str()});
  }
  void param(const std::string& region, const std::string& cat, const std::string& v) {
    params_.push_back({region, cat, "SINGLE", v});
  }
  void print(const std::string& file) const {
    if (file.empty()) {
      write(std::cout);
      return;
    }
    std::ofstream out(file);
    if (!out.good()) {
      std::cout << "WARNING: Could not open stats file for writing, file provided:" << file << "\n";
      write(std::cerr);
      return;
    }
    write(out);
  }

 private:
  struct Row {
    std::string region, cat, total, val;
  };
  void write(std::ostream& out) const {
    const char* SEP = ", ";
    out << "STAT_TYPE" << SEP << "REGION" << SEP << "CATEGORY" << SEP << "TOTAL_TYPE" << SEP << "TOTAL\n";
    for (const auto& r : ints_) out << "STAT" << SEP << r.region << SEP << r.cat << SEP << r.total << SEP << r.val << "\n";
    for (const auto& r : fps_) out << "STAT" << SEP << r.region << SEP << r.cat << SEP << r.total << SEP << r.val << "\n";
    for (const auto& r : params_)
      out << "PARAM" << SEP << r.region << SEP << r.cat << SEP << r.total << SEP << r.val << "\n";
  }
  std::vector<Row> ints_, fps_, params_;
};

// Wall-clock timer reported like galois::StatTimer (milliseconds, TMAX).
class Timer {
 public:
  Timer(Stats& s, const char* name, const char* region = "(NULL)") : s_(s), name_(name), region_(region) {}
  void start() { t0_ = std::chrono::steady_clock::now(); }
  void stop() {
    ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
  }
  void report() { s_.stat_int(region_, name_, "TMAX", static_cast<int64_t>(ms_)); }
  double ms() const { return ms_; }

 private:
  Stats& s_;
  const char* name_;
  const char* region_;
  std::chrono::steady_clock::time_point t0_{};
  double ms_ = 0;
};

// Banner and the standard parameters of LonestarStart.
inline void start(const Cli& cli, Stats& stats, const char* app, const char* desc) {
  std::cout << "LoneStar Benchmark Suite (galois_amd, HIP gfx950)\n";
  std::cout << "application: " << app << "\n";
  if (desc) std::cout << desc << "\n";
  std::cout << "\n";
  std::cout.flush();
  stats.param("(NULL)", "CommandLine", cli.cmdline());
  stats.param("(NULL)", "Threads", cli.str("t"));
  stats.param("(NULL)", "Hosts", "1");
  stats.param("(NULL)", "Input", cli.input());
  char host[256] = {0};
  gethostname(host, sizeof(host) - 1);
  stats.param("(NULL)", "Hostname", host);
  check(gx_init(0), "gx_init");
}

// ------------------------------------------------------------ graph input
struct HostGraph {
  gx_host_csr csr{};
  ~HostGraph() { gx_host_csr_free(&csr); }
  int64_t n() const { return csr.n; }
  int64_t m() const { return csr.m; }
};

inline void read_graph(const std::string& path, HostGraph& hg) {
  check(gx_gr_read(path.c_str(), &hg.csr), "reading the input graph");
}

// Device graph built from the host CSR; weights attached when the file
// carries 4-byte edge data and `weighted` is asked.
inline gx_graph* upload(const HostGraph& hg, bool symmetric, bool weighted) {
  uint32_t flags = symmetric ? GX_GRAPH_SYMMETRIC : 0;
  const void* w = nullptr;
  if (weighted) {
    if (hg.csr.edge_size != 4 || !hg.csr.edge_data) die("this application needs 4-byte (uint32) edge weights");
    flags |= GX_GRAPH_WEIGHTED;
    w = hg.csr.edge_data;
  }
  gx_graph* g = nullptr;
  check(gx_graph_create(hg.n(), hg.m(), hg.csr.row, hg.csr.col, w, flags, &g), "gx_graph_create");
  return g;
}

template <class T>
struct DeviceArray {
  T* p = nullptr;
  size_t n = 0;
  explicit DeviceArray(size_t count) : n(count) {
    hip_check(hipMalloc(&p, (count ? count : 1) * sizeof(T)), "hipMalloc");
  }
  ~DeviceArray() { (void)hipFree(p); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  std::vector<T> host() const {
    std::vector<T> h(n);
    if (n) hip_check(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
  }
};

// isApproximateDegreeDistributionPowerLaw of lonestar/liblonestar/include/
// Lonestar/Utils.h:37-57 through the C ABI (gx_degree_power_law).
inline bool approximately_power_law(const HostGraph& hg) {
  int power_law = 0;
  check(gx_degree_power_law(hg.n(), hg.m(), hg.csr.row, &power_law), "gx_degree_power_law");
  return power_law != 0;
}

// Per-node output of the LonestarGPU apps (the skelapp -o option,
// libgpu/src/skelapp/skel.cu:75,105; format of lonestar/analytics/gpu/bfs/
// support.cu:27-47): one "node value" line per node, read by
// scripts/result_checker.py.  `fmt` prints one value.
template <class T, class F>
void write_node_values(const std::string& path, const std::vector<T>& values, F fmt) {
  if (path.empty()) return;
  FILE* f = path == "-" ? stdout : std::fopen(path.c_str(), "w");
  if (!f) die("cannot open output file " + path);
  for (size_t i = 0; i < values.size(); ++i) {
    if (std::fprintf(f, "%zu ", i) < 0 || !fmt(f, values[i])) die("write to " + path + " failed");
  }
  if (f != stdout && std::fclose(f) != 0) die("write to " + path + " failed");
  if (f == stdout) std::fflush(stdout);
}

// Distances and levels: unreached vertices are printed as UINT32_MAX / 4
// (1073741823), the value of the distributed apps' output, as
// lonestar/analytics/gpu/bfs/support.cu:38-45 and sssp/support.cu:45-52 do.
inline void write_distances(const std::string& path, const std::vector<uint32_t>& d) {
  write_node_values(path, d, [](FILE* f, uint32_t v) {
    return std::fprintf(f, "%u\n", v == GX_DIST_INFINITY ? 1073741823u : v) > 0;
  });
}

// The 20 largest ranks, ties to the lower id (TopPair ordering and printTop
// of lonestar/analytics/cpu/pagerank/PageRank-constants.h:60-105).
inline void print_top(const std::vector<float>& x, unsigned topn = 20) {
  struct Key {
    float value;
    uint32_t id;
    bool operator<(const Key& b) const { return value == b.value ? id > b.id : value < b.value; }
  };
  std::map<Key, uint32_t> top;
  for (uint32_t v = 0; v < x.size(); ++v) {
    const Key k{x[v], v};
    if (top.size() < topn) {
      top.emplace(k, v);
      continue;
    }
    if (top.begin()->first < k) {
      top.erase(top.begin());
      top.emplace(k, v);
    }
  }
  std::cout << "Rank PageRank Id\n";
  int rank = 1;
  for (auto it = top.rbegin(); it != top.rend(); ++it, ++rank)
    std::cout << rank << ": " << it->first.value << " " << it->first.id << "\n";
}

// Ranks per node, "%d %f"-style lines as lonestar/analytics/gpu/pagerank/
// support.cu:95-97, with 9 significant digits (the %f of the reference
// prints normalised ranks as 0.000000).
inline void write_ranks(const std::string& path, const std::vector<float>& x) {
  write_node_values(path, x, [](FILE* f, float v) { return std::fprintf(f, "%.9g\n", v) > 0; });
}

}  // namespace lonestar
