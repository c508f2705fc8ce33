// graph-convert: the subset of the reference's converter that the lonestar
// analytics inputs go through (tools/graph-convert/graph-convert.cpp):
//   -edgelist2gr  -csv2gr       text edge lists -> .gr (convertEdgelist
//                               :351-529: one "src dst [value]" per line,
//                               lines that do not parse are skipped, n = max
//                               id + 1, neighbours kept in input order; csv
//                               skips a header line and uses ',')
//   -gr2edgelist  -gr2edgelist1ind   .gr -> "src dst [value]" text (:867-925)
//   -gr2sgr       symmetrise by adding every reverse edge except for
//                 self-loops (libgalois/include/galois/graphs/FileGraph.h:
//                 764-825)
//   -gr2tgr       transpose, in-edges in ascending source order (:1831-1880)
//   -gr2cgr       remove self-loops and duplicate edges, keeping the first of
//                 each duplicate after sorting by destination (:1902-1994)
//   -gr2randomweightgr  edge values drawn by std::mt19937 (default seed) and
//                 std::uniform_int_distribution / uniform_real_distribution
//                 over [-minValue, -maxValue] (:1145-1192)
// -edgeType selects the edge value type (void, int32, int64, uint32, uint64,
// float32, float64).  Host only: no GPU is used.
#include <array>
#include <cstring>
#include <random>

#include "lonestar.hpp"

using namespace lonestar;

namespace {

enum class Ty { Void, I32, I64, U32, U64, F32, F64 };

struct TyInfo {
  const char* name;
  Ty ty;
  uint32_t size;
};
const TyInfo kTypes[] = {{"void", Ty::Void, 0},  {"int32", Ty::I32, 4},   {"int64", Ty::I64, 8}, {"uint32", Ty::U32, 4},
                         {"uint64", Ty::U64, 8}, {"float32", Ty::F32, 4}, {"float64", Ty::F64, 8}};

const TyInfo& type_of(const std::string& s) {
  for (const auto& t : kTypes)
    if (s == t.name) return t;
  die("unknown -edgeType " + s);
}

// A CSR with raw edge values of `esize` bytes per edge.
struct Csr {
  int64_t n = 0;
  std::vector<int32_t> row{0}, col;
  std::vector<uint8_t> data;
  uint32_t esize = 0;
  int64_t m() const { return static_cast<int64_t>(col.size()); }
};

Csr load(const std::string& path) {
  HostGraph hg;
  read_graph(path, hg);
  Csr g;
  g.n = hg.n();
  g.row.assign(hg.csr.row, hg.csr.row + hg.n() + 1);
  g.col.assign(hg.csr.col, hg.csr.col + hg.m());
  g.esize = hg.csr.edge_size;
  if (g.esize && hg.csr.edge_data) {
    const uint8_t* p = static_cast<const uint8_t*>(hg.csr.edge_data);
    g.data.assign(p, p + static_cast<size_t>(hg.m()) * g.esize);
  }
  return g;
}

void save(const std::string& path, const Csr& g) {
  if (g.n > INT32_MAX || g.m() > INT32_MAX) die("graph too large for int32 offsets");
  gx_host_csr h{};
  h.n = g.n;
  h.m = g.m();
  h.row = const_cast<int32_t*>(g.row.data());
  h.col = const_cast<int32_t*>(g.col.data());
  h.edge_data = g.esize ? const_cast<uint8_t*>(g.data.data()) : nullptr;
  h.edge_size = g.esize;
  check(gx_gr_write(path.c_str(), &h), "writing the output graph");
}

void status(int64_t in_n, int64_t in_m, int64_t out_n, int64_t out_m) {
  std::cout << "InGraph : |V| = " << in_n << ", |E| = " << in_m << "\n";
  std::cout << "OutGraph: |V| = " << out_n << ", |E| = " << out_m << "\n";
}

// Parses one value of type t from the stream into 8 bytes of raw storage.
bool parse_value(std::istream& is, const TyInfo& t, uint8_t* out) {
  switch (t.ty) {
    case Ty::I32: { int32_t v; if (!(is >> v)) return false; std::memcpy(out, &v, 4); return true; }
    case Ty::I64: { int64_t v; if (!(is >> v)) return false; std::memcpy(out, &v, 8); return true; }
    case Ty::U32: { uint32_t v; if (!(is >> v)) return false; std::memcpy(out, &v, 4); return true; }
    case Ty::U64: { uint64_t v; if (!(is >> v)) return false; std::memcpy(out, &v, 8); return true; }
    case Ty::F32: { float v; if (!(is >> v)) return false; std::memcpy(out, &v, 4); return true; }
    case Ty::F64: { double v; if (!(is >> v)) return false; std::memcpy(out, &v, 8); return true; }
    default: return true;
  }
}

void print_value(std::ostream& os, const TyInfo& t, const uint8_t* p) {
  switch (t.ty) {
    case Ty::I32: { int32_t v; std::memcpy(&v, p, 4); os << v; break; }
    case Ty::I64: { int64_t v; std::memcpy(&v, p, 8); os << v; break; }
    case Ty::U32: { uint32_t v; std::memcpy(&v, p, 4); os << v; break; }
    case Ty::U64: { uint64_t v; std::memcpy(&v, p, 8); os << v; break; }
    case Ty::F32: { float v; std::memcpy(&v, p, 4); os << v; break; }
    case Ty::F64: { double v; std::memcpy(&v, p, 8); os << v; break; }
    default: break;
  }
}

struct Edge {
  uint64_t src, dst;
  uint8_t val[8];
};

// convertEdgelist: both passes of the reference over the same lines.
Csr edgelist(const std::string& path, const TyInfo& t, bool skip_first, bool csv) {
  std::ifstream in(path);
  if (!in) die("failed to open input file " + path);
  std::vector<Edge> edges;
  std::string line;
  size_t lineno = 0;
  bool skipped = false;
  size_t skipped_line = 0;
  if (skip_first) {
    std::cout << "WARNING: first line is assumed to contain labels and will be ignored\n";
    std::getline(in, line);
    ++lineno;
  }
  for (; std::getline(in, line); ++lineno) {
    std::stringstream is(line);
    Edge e{};
    char d;
    bool ok = static_cast<bool>(is >> e.src);
    if (ok && csv) ok = (is >> d) && d == ',';
    if (ok) ok = static_cast<bool>(is >> e.dst);
    if (ok && t.ty != Ty::Void) {
      if (csv) ok = (is >> d) && d == ',';
      if (ok) ok = parse_value(is, t, e.val);
    }
    if (!ok) {
      skipped = true;
      skipped_line = lineno;
      continue;
    }
    edges.push_back(e);
  }
  if (skipped)
    std::cout << "WARNING: ignored at least one line (line " << skipped_line
              << ") because it did not match the expected format\n";
  uint64_t maxid = 0;
  for (const Edge& e : edges) maxid = std::max({maxid, e.src, e.dst});
  Csr g;
  g.n = static_cast<int64_t>(maxid) + 1;
  g.esize = t.size;
  g.row.assign(g.n + 1, 0);
  for (const Edge& e : edges) ++g.row[e.src + 1];
  for (int64_t v = 0; v < g.n; ++v) g.row[v + 1] += g.row[v];
  g.col.resize(edges.size());
  g.data.resize(edges.size() * g.esize);
  std::vector<int32_t> cur(g.row.begin(), g.row.end() - 1);
  for (const Edge& e : edges) {
    const int32_t at = cur[e.src]++;
    g.col[at] = static_cast<int32_t>(e.dst);
    if (g.esize) std::memcpy(&g.data[static_cast<size_t>(at) * g.esize], e.val, g.esize);
  }
  status(g.n, g.m(), g.n, g.m());
  return g;
}

void to_edgelist(const Csr& g, const TyInfo& t, const std::string& out, int base) {
  std::ofstream f(out);
  const bool vals = t.ty != Ty::Void && g.esize == t.size && g.esize;
  for (int64_t u = 0; u < g.n; ++u)
    for (int32_t e = g.row[u]; e < g.row[u + 1]; ++e) {
      f << u + base << " " << g.col[e] + base;
      if (vals) {
        f << " ";
        print_value(f, t, &g.data[static_cast<size_t>(e) * g.esize]);
      }
      f << "\n";
    }
  status(g.n, g.m(), g.n, g.m());
}

// Builds a CSR from (src, dst, edge index) triples kept in the given order.
Csr from_pairs(int64_t n, const std::vector<std::array<int64_t, 3>>& es, const Csr& in) {
  Csr g;
  g.n = n;
  g.esize = in.esize;
  g.row.assign(n + 1, 0);
  for (const auto& e : es) ++g.row[e[0] + 1];
  for (int64_t v = 0; v < n; ++v) g.row[v + 1] += g.row[v];
  g.col.resize(es.size());
  g.data.resize(es.size() * g.esize);
  std::vector<int32_t> cur(g.row.begin(), g.row.end() - 1);
  for (const auto& e : es) {
    const int32_t at = cur[e[0]]++;
    g.col[at] = static_cast<int32_t>(e[1]);
    if (g.esize)
      std::memcpy(&g.data[static_cast<size_t>(at) * g.esize], &in.data[static_cast<size_t>(e[2]) * g.esize],
                  g.esize);
  }
  return g;
}

Csr symmetrize(const Csr& in) {
  std::vector<std::array<int64_t, 3>> es;
  es.reserve(in.col.size() * 2);
  for (int64_t u = 0; u < in.n; ++u)
    for (int32_t e = in.row[u]; e < in.row[u + 1]; ++e) {
      es.push_back({u, in.col[e], e});
      if (in.col[e] != u) es.push_back({in.col[e], u, e});
    }
  Csr g = from_pairs(in.n, es, in);
  status(in.n, in.m(), g.n, g.m());
  return g;
}

Csr transpose(const Csr& in) {
  std::vector<std::array<int64_t, 3>> es;
  es.reserve(in.col.size());
  for (int64_t u = 0; u < in.n; ++u)
    for (int32_t e = in.row[u]; e < in.row[u + 1]; ++e) es.push_back({in.col[e], u, e});
  Csr g = from_pairs(in.n, es, in);
  status(in.n, in.m(), g.n, g.m());
  return g;
}

Csr cleanup(const Csr& in, bool& unchanged) {
  std::vector<std::array<int64_t, 3>> es;
  for (int64_t u = 0; u < in.n; ++u) {
    std::vector<std::pair<int32_t, int32_t>> adj;  // (dst, edge)
    for (int32_t e = in.row[u]; e < in.row[u + 1]; ++e) adj.push_back({in.col[e], e});
    std::stable_sort(adj.begin(), adj.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t i = 0; i < adj.size(); ++i) {
      if (adj[i].first == u) continue;
      if (i > 0 && adj[i - 1].first == adj[i].first) continue;
      es.push_back({u, adj[i].first, adj[i].second});
    }
  }
  unchanged = static_cast<int64_t>(es.size()) == in.m();
  Csr g = from_pairs(in.n, es, in);
  if (unchanged) {
    std::cout << "Graph already simplified; copy input to output\n";
    status(in.n, in.m(), in.n, in.m());
  } else {
    status(in.n, in.m(), g.n, g.m());
  }
  return g;
}

template <class T>
void fill_random(Csr& g, int lo, int hi) {
  std::mt19937 gen;
  std::vector<uint8_t> d(g.col.size() * sizeof(T));
  if constexpr (std::is_integral<T>::value) {
    std::uniform_int_distribution<T> dist(lo, hi);
    for (size_t e = 0; e < g.col.size(); ++e) {
      const T v = dist(gen);
      std::memcpy(&d[e * sizeof(T)], &v, sizeof(T));
    }
  } else {
    std::uniform_real_distribution<T> dist(lo, hi);
    for (size_t e = 0; e < g.col.size(); ++e) {
      const T v = dist(gen);
      std::memcpy(&d[e * sizeof(T)], &v, sizeof(T));
    }
  }
  g.data.swap(d);
  g.esize = sizeof(T);
}

}  // namespace

int main(int argc, char** argv) {
  const char* kinds[] = {"edgelist2gr", "csv2gr", "gr2edgelist", "gr2edgelist1ind", "gr2sgr",
                         "gr2tgr",      "gr2cgr", "gr2randomweightgr"};
  // The two positional file names are taken before the generic parser sees
  // the options.
  std::vector<char*> args;
  std::vector<std::string> files;
  std::string kind;
  for (int i = 0; i < argc; ++i) {
    const std::string a = argv[i];
    bool is_kind = false;
    for (const char* k : kinds)
      if (a == std::string("-") + k || a == std::string("--") + k) {
        kind = k;
        is_kind = true;
      }
    if (is_kind) continue;
    if (i > 0 && !a.empty() && a[0] != '-') {
      files.push_back(a);
      continue;
    }
    args.push_back(argv[i]);
  }
  Cli cli;
  cli.value("edgeType", "void", "Input/Output edge type: void, int32, int64, uint32, uint64, float32, float64");
  cli.value("minValue", "1", "minimum weight to add for random weight conversions");
  cli.value("maxValue", "100", "maximum weight to add for random weight conversions");
  if (files.size() == 2) args.push_back(const_cast<char*>(files[0].c_str()));
  cli.parse(static_cast<int>(args.size()), args.data(), "graph-convert");
  if (kind.empty()) die("graph-convert: choose a conversion (-edgelist2gr, -csv2gr, -gr2edgelist, ...)");
  if (files.size() != 2) die("graph-convert: expects <input file> <output file>");
  const TyInfo& t = type_of(cli.str("edgeType"));
  std::cout << "Graph type: " << t.name << "\n";
  const std::string& in = files[0];
  const std::string& out = files[1];
  if (kind == "edgelist2gr" || kind == "csv2gr") {
    save(out, edgelist(in, t, kind == "csv2gr", kind == "csv2gr"));
  } else if (kind == "gr2edgelist" || kind == "gr2edgelist1ind") {
    to_edgelist(load(in), t, out, kind == "gr2edgelist1ind" ? 1 : 0);
  } else if (kind == "gr2sgr") {
    save(out, symmetrize(load(in)));
  } else if (kind == "gr2tgr") {
    save(out, transpose(load(in)));
  } else if (kind == "gr2cgr") {
    bool unchanged = false;
    const Csr g = load(in);
    const Csr c = cleanup(g, unchanged);
    save(out, unchanged ? g : c);
  } else if (kind == "gr2randomweightgr") {
    if (t.ty == Ty::Void) die("gr2randomweightgr needs an -edgeType with values");
    Csr g = load(in);
    const int lo = static_cast<int>(cli.f64("minValue")), hi = static_cast<int>(cli.f64("maxValue"));
    switch (t.ty) {
      case Ty::I32: fill_random<int32_t>(g, lo, hi); break;
      case Ty::I64: fill_random<int64_t>(g, lo, hi); break;
      case Ty::U32: fill_random<uint32_t>(g, lo, hi); break;
      case Ty::U64: fill_random<uint64_t>(g, lo, hi); break;
      case Ty::F32: fill_random<float>(g, lo, hi); break;
      case Ty::F64: fill_random<double>(g, lo, hi); break;
      default: break;
    }
    save(out, g);
    status(g.n, g.m(), g.n, g.m());
  }
  return 0;
}
