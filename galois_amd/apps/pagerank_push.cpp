// pagerank-push: the lonestar push PageRank app entry point (lonestar/
// analytics/cpu/pagerank/PageRank-push.cpp: -algo :43-48, main :188-250,
// flags and printTop PageRank-constants.h:32-52,78-105) on the GPU residual
// push of the C ABI (gx_pagerank_push: value starts 0, residual 1 - alpha,
// a vertex whose residual exceeds the tolerance moves it into its value and
// pushes alpha * residual / outdeg to its out-neighbours; unnormalised, as
// the reference).  The input is the graph itself (out-edges).  -algo=Async
// (the reference's default) and -algo=Sync both run the synchronous rounds
// of syncPageRank (:103-188); the asynchronous schedule only changes the
// order of the FP32 additions.
#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("tolerance", "0.001", "tolerance");
  cli.value("maxIterations", "1000", "Maximum iterations, applies round-based versions only");
  cli.value("algo", "Async", "Choose an algorithm: Async, Sync");
  cli.parse(argc, argv, "pagerank-push");
  Stats stats;
  start(cli, stats, "Page Rank", "Computes page ranks a la Page and Brin. This is a push-style algorithm.");
  Timer total(stats, "TimerTotal");
  total.start();
  HostGraph hg;
  read_graph(cli.input(), hg);
  std::cout << "Read " << hg.n() << " nodes, " << hg.m() << " edges\n";
  const float tol = static_cast<float>(cli.f64("tolerance"));
  const uint32_t maxit = static_cast<uint32_t>(cli.u64("maxIterations"));
  std::cout << "tolerance:" << tol << ", maxIterations:" << maxit << "\n";
  const std::string algo = cli.str("algo");
  if (algo != "Async" && algo != "Sync") die("unknown -algo " + algo);
  gx_graph* g = upload(hg, cli.on("symmetricGraph"), false);
  DeviceArray<float> pr(hg.n());
  Timer exec(stats, "Timer_0");
  exec.start();
  std::cout << "Running Edge " << algo << " push version,";
  gx_stats st{};
  check(gx_pagerank_push(g, 0.85f, tol, maxit, pr.p, &st), "gx_pagerank_push");
  hip_check(hipDeviceSynchronize(), "pagerank");
  exec.stop();
  std::cout << " iterations:" << st.iterations << "\n";
  if (st.iterations >= maxit) std::cerr << "ERROR: failed to converge in " << st.iterations << " iterations\n";
  const std::vector<float> x = pr.host();
  write_ranks(cli.str("o"), x);
  if (!cli.on("noverify")) print_top(x);
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
