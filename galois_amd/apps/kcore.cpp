// k-core: the lonestar k-core app entry point (lonestar/analytics/cpu/k-core/
// kcore.cpp: flags :40-55, kCoreSanity :184-202, main :208-290) on the GPU
// cascade peeling of the C ABI.  The k-core is unique, so both -algo choices
// report the same count.
#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("algo", "Sync", "Choose an algorithm (default Sync): Async, Sync");
  cli.value("kcore", "", "k-core value");
  cli.parse(argc, argv, "k-core");
  if (!cli.given("kcore")) die("k-core: option 'kcore' must be specified at least once!");
  Stats stats;
  start(cli, stats, "k-core", "Finds the k-core of a graph, defined as the subgraph where all vertices have degree at "
                               "least k.");
  Timer total(stats, "TimerTotal");
  total.start();
  if (!cli.on("symmetricGraph"))
    die("This application requires a symmetric graph input; please use the -symmetricGraph flag  to indicate "
        "the input is a symmetric graph.");
  const uint32_t k = static_cast<uint32_t>(cli.u64("kcore"));
  HostGraph hg;
  Timer readt(stats, "GraphConstructTime", "KCore");
  readt.start();
  read_graph(cli.input(), hg);
  gx_graph* g = upload(hg, true, false);
  readt.stop();
  readt.report();
  DeviceArray<uint32_t> degree(hg.n());
  Timer exec(stats, "Timer_0");
  exec.start();
  if (cli.str("algo") == "Async")
    info("Running asynchronous k-core with k-core number ", k);
  else
    info("Running synchronous k-core with k-core number ", k);
  uint64_t alive = 0;
  check(gx_kcore(g, k, degree.p, nullptr, &alive, nullptr), "gx_kcore");
  hip_check(hipDeviceSynchronize(), "kcore");
  exec.stop();
  if (!cli.on("noverify")) print("Number of nodes in the ", k, "-core is ", alive, "\n");
  // "node flag" lines: 1 for members of the k-core (the flag field of
  // lonestar/analytics/distributed/k-core/kcore_push.cpp:69).
  if (!cli.str("o").empty()) {
    const std::vector<uint32_t> deg = degree.host();
    write_node_values(cli.str("o"), deg, [k](FILE* f, uint32_t d) { return std::fprintf(f, "%d\n", d >= k) > 0; });
  }
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
