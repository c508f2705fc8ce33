// pagerank-pull: the lonestar pull PageRank app entry point (lonestar/
// analytics/cpu/pagerank/PageRank-pull.cpp: flags :34-44 and
// PageRank-constants.h:32-52, main :276-357, printTop PageRank-constants.h:
// 78-105) with the reference's own unnormalised semantics on the GPU
// (gx_pagerank_pull_ref: -algo=Residual is the two-sweep residual pull,
// bit-exact with the reference for the same in-edge order; -algo=Topo is the
// Jacobi topological pull).  The input is the transposed graph, as in the
// reference.
#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("tolerance", "0.001", "tolerance");
  cli.value("maxIterations", "1000", "Maximum iterations, applies round-based versions only");
  cli.value("algo", "Residual", "Choose an algorithm: Topo, Residual");
  cli.flag("transposedGraph", "Specify that the input graph is transposed");
  cli.parse(argc, argv, "pagerank-pull");
  Stats stats;
  start(cli, stats, "Page Rank", "Computes page ranks a la Page and Brin. This is a pull-style algorithm.");
  if (!cli.on("transposedGraph"))
    die("This application requires a transposed graph input; please use the -transposedGraph flag  to indicate "
        "the input is a transposed graph.");
  Timer total(stats, "TimerTotal");
  total.start();
  std::cout << "WARNING: pull style algorithms work on the transpose of the actual graph\n"
            << "WARNING: this program assumes that " << cli.input() << " contains transposed representation\n\n"
            << "Reading graph: " << cli.input() << "\n";
  HostGraph hg;
  read_graph(cli.input(), hg);
  std::cout << "Read " << hg.n() << " nodes, " << hg.m() << " edges\n";
  const float tol = static_cast<float>(cli.f64("tolerance"));
  const uint32_t maxit = static_cast<uint32_t>(cli.u64("maxIterations"));
  const std::string algo = cli.str("algo");
  int variant = 0;
  if (algo == "Topo") {
    variant = 1;
    std::cout << "Running Pull Topological version, tolerance:" << tol << ", maxIterations:" << maxit << "\n";
  } else if (algo == "Residual") {
    std::cout << "Running Pull Residual version, tolerance:" << tol << ", maxIterations:" << maxit << "\n";
  } else {
    die("unknown -algo " + algo);
  }
  // The file holds the transpose; its transpose is the graph itself.
  gx_graph* gt = upload(hg, false, false);
  check(gx_graph_build_transpose(gt), "gx_graph_build_transpose");
  const int32_t *trow = nullptr, *tcol = nullptr;
  check(gx_graph_device_arrays(gt, nullptr, nullptr, nullptr, &trow, &tcol, nullptr), "gx_graph_device_arrays");
  gx_graph* g = nullptr;
  check(gx_graph_create_device(hg.n(), hg.m(), const_cast<int32_t*>(trow), const_cast<int32_t*>(tcol), nullptr, 0,
                               &g),
        "gx_graph_create_device");
  DeviceArray<float> pr(hg.n());
  Timer exec(stats, "Timer_0");
  exec.start();
  check(gx_pagerank_pull_ref(g, variant, 0.85f, tol, maxit, pr.p, nullptr), "gx_pagerank_pull_ref");
  hip_check(hipDeviceSynchronize(), "pagerank");
  exec.stop();
  const std::vector<float> x = pr.host();
  float mx = x.empty() ? 0.f : x[0], mn = mx, sum = 0.f;
  for (float v : x) {
    mx = std::max(mx, v);
    mn = std::min(mn, v);
    sum += v;
  }
  info("Max rank is ", mx);
  info("Min rank is ", mn);
  info("Sum is ", sum);
  write_ranks(cli.str("o"), x);
  if (!cli.on("noverify")) print_top(x);
  check(gx_graph_free(g), "gx_graph_free");
  check(gx_graph_free(gt), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
