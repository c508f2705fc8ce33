// betweennesscentrality: the lonestar BC app entry point (lonestar/analytics/
// cpu/betweennesscentrality/BetweennessCentrality.cpp: flags :41-81, main
// :99-145; source selection and output of LevelStructs.h:224-326) on the GPU
// Brandes kernels of the C ABI (sources batched 64 per wavefront when there
// are many, one per pass otherwise).  bc is summed in FP64 over the sources
// and printed as the reference's float ("%u %.9f" with -output).
#include <iterator>

#include "lonestar.hpp"

using namespace lonestar;

int main(int argc, char** argv) {
  Cli cli;
  cli.value("sourcesToUse", "", "Whitespace separated list of sources in a file to use in BC");
  cli.value("numOfSources", "0", "Number of sources to compute BC on (default all)");
  cli.value("numOfOutSources", "0", "Number of sources WITH EDGES to compute BC on (default is all)");
  cli.flag("singleSource", "Level: Use for single source BC (default off)");
  cli.value("startNode", "0", "Level/Outer: Starting source node used for betweeness-centrality (default 0)");
  cli.flag("output", "Output BC (Level/Async) (default: false)");
  cli.value("algo", "AutoAlgo", "Choose an algorithm (default value AutoAlgo): Level, Async, Outer, AutoAlgo");
  cli.parse(argc, argv, "betweennesscentrality");
  Stats stats;
  start(cli, stats, "Betweenness Centrality", "Computes betwenness centrality in an unweighted graph");
  Timer total(stats, "TimerTotal");
  total.start();
  HostGraph hg;
  read_graph(cli.input(), hg);
  std::string algo = cli.str("algo");
  if (algo == "AutoAlgo") {
    algo = approximately_power_law(hg) ? "Async" : "Level";
    info("Choosing ", algo, " algorithm");
  }
  if (algo == "Level")
    info("Running level BC");
  else if (algo == "Async")
    info("Running async BC");
  else if (algo == "Outer")
    info("Running outer BC");
  else
    die("Unknown BC algorithm type");
  gx_graph* g = upload(hg, cli.on("symmetricGraph"), false);
  info("Graph construction complete");

  std::vector<uint32_t> file_sources;
  if (!cli.str("sourcesToUse").empty()) {
    std::ifstream f(cli.str("sourcesToUse"));
    file_sources.assign(std::istream_iterator<uint64_t>{f}, std::istream_iterator<uint64_t>{});
  }
  std::vector<uint32_t> sources;
  if (cli.on("singleSource")) {
    sources.push_back(static_cast<uint32_t>(cli.u64("startNode")));
  } else {
    uint64_t loop_end = cli.u64("numOfSources") ? cli.u64("numOfSources") : static_cast<uint64_t>(hg.n());
    if (!file_sources.empty()) {
      loop_end = std::min<uint64_t>(loop_end, file_sources.size());
      sources.assign(file_sources.begin(), file_sources.begin() + loop_end);
    } else {
      for (uint64_t i = 0; i < loop_end; ++i) sources.push_back(static_cast<uint32_t>(i));
    }
  }
  for (uint32_t s : sources)
    if (static_cast<int64_t>(s) >= hg.n()) die("source " + std::to_string(s) + " out of range");

  DeviceArray<double> bc(hg.n());
  info("Beginning main computation");
  Timer exec(stats, "Timer_0");
  exec.start();
  if (sources.size() >= 64)
    check(gx_bc_batched(g, sources.data(), static_cast<uint32_t>(sources.size()), bc.p, nullptr), "gx_bc_batched");
  else
    check(gx_bc(g, sources.data(), static_cast<uint32_t>(sources.size()), bc.p, nullptr), "gx_bc");
  hip_check(hipDeviceSynchronize(), "bc");
  exec.stop();
  const std::vector<double> b = bc.host();
  float mx = b.empty() ? 0.f : static_cast<float>(b[0]), mn = mx, sum = 0.f;
  for (double v : b) {
    mx = std::max(mx, static_cast<float>(v));
    mn = std::min(mn, static_cast<float>(v));
    sum += static_cast<float>(v);
  }
  print("Max BC is ", mx, "\n");
  print("Min BC is ", mn, "\n");
  print("BC sum is ", sum, "\n");
  if (cli.on("output")) {
    char line[64];
    for (int64_t v = 0; v < hg.n(); ++v) {
      std::snprintf(line, sizeof(line), "%u %.9f\n", static_cast<unsigned>(v), static_cast<float>(b[v]));
      std::cout << line;
    }
  }
  write_node_values(cli.str("o"), b,
                    [](FILE* f, double v) { return std::fprintf(f, "%.9f\n", static_cast<float>(v)) > 0; });
  check(gx_graph_free(g), "gx_graph_free");
  total.stop();
  exec.report();
  total.report();
  stats.print(cli.str("statFile"));
  return 0;
}
