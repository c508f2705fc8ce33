// Layout and work tables of the propagation-blocked PageRank sweep
// (pr_blocking.hpp), built once per graph from the CSR and the degree order.
//
//  1. count_kernel: edges per (chunk of the source, bin of the destination)
//     key in degree-order ids, one atomic per run of equal keys in a
//     wavefront.  The edges are read in out-edge (CSR) order, so that the
//     edges of one source stay together and neighbouring entries of a run
//     have distinct destinations: phase 2 then adds a wavefront's 64 entries
//     into mostly different LDS words (destination order would put up to 64
//     adds on one word) and phase 1 reads one source's y as a broadcast.
//  2. exclusive scan of the counts: start of every run, chunk-major.
//  3. scatter_kernel: every run of equal keys in a wavefront reserves its
//     slots with one atomic and writes src16/dst16.  The order of the edges
//     inside a run depends on the atomics; the sweep's integer sums do not.
//  4. host: phase-1 pieces per chunk, phase-2 jobs per bin in three length
//     classes, tasks (bins, heavy bins sliced), sorted longest first.
#include <algorithm>
#include <vector>

#include "pr_blocking.hpp"

namespace gx {
namespace pb {
namespace {

constexpr int kBuildBlock = 256;

// Per wavefront: lanes whose key differs from the previous lane's start a
// run; returns the run heads as a mask and the length of the caller's run
// (valid on head lanes).
__device__ __forceinline__ uint64_t run_heads(bool valid, uint32_t key, unsigned lane, unsigned nvalid,
                                              unsigned& len) {
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = valid && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(head);
  const uint64_t above = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
  const unsigned next = above ? static_cast<unsigned>(__builtin_ctzll(above)) : nvalid;
  len = next - lane;
  return heads;
}

__global__ void __launch_bounds__(kBuildBlock) count_kernel(const uint32_t* __restrict__ esrc,
                                                            const int32_t* __restrict__ col,
                                                            const uint32_t* __restrict__ rank, int64_t m,
                                                            uint32_t nb, uint32_t* __restrict__ cnt) {
  const unsigned lane = lane_id();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e0 = blockIdx.x * static_cast<int64_t>(kBuildBlock) + (threadIdx.x & ~(kWave - 1)); e0 < m;
       e0 += stride) {
    const int64_t e = e0 + lane;
    const bool valid = e < m;
    const unsigned nvalid = static_cast<unsigned>(m - e0 < kWave ? m - e0 : kWave);
    const uint32_t key =
        valid ? (rank[esrc[e]] >> kShift) * nb + (rank[col[e]] >> kShift) : 0xffffffffu;
    unsigned len;
    const uint64_t heads = run_heads(valid, key, lane, nvalid, len);
    if ((heads >> lane) & 1ull) atomicAdd(&cnt[key], len);
  }
}

__global__ void __launch_bounds__(kBuildBlock) scatter_kernel(const uint32_t* __restrict__ esrc,
                                                              const int32_t* __restrict__ col,
                                                              const uint32_t* __restrict__ rank, int64_t m,
                                                              uint32_t nb, uint32_t* __restrict__ cursor,
                                                              uint16_t* __restrict__ src16,
                                                              uint16_t* __restrict__ dst16) {
  const unsigned lane = lane_id();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e0 = blockIdx.x * static_cast<int64_t>(kBuildBlock) + (threadIdx.x & ~(kWave - 1)); e0 < m;
       e0 += stride) {
    const int64_t e = e0 + lane;
    const bool valid = e < m;
    const unsigned nvalid = static_cast<unsigned>(m - e0 < kWave ? m - e0 : kWave);
    const uint32_t u = valid ? rank[esrc[e]] : 0u;
    const uint32_t v = valid ? rank[col[e]] : 0u;
    const uint32_t key = valid ? (u >> kShift) * nb + (v >> kShift) : 0xffffffffu;
    unsigned len;
    const uint64_t heads = run_heads(valid, key, lane, nvalid, len);
    uint32_t base = 0;
    if ((heads >> lane) & 1ull) base = atomicAdd(&cursor[key], len);
    // Head of the caller's run: highest head at or below the lane.
    const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const unsigned h = 63u - static_cast<unsigned>(__builtin_clzll((heads & le) | 1ull));
    base = __shfl(base, static_cast<int>(h));
    if (valid) {
      const uint32_t pos = base + (lane - h);
      src16[pos] = static_cast<uint16_t>(u & (kSpan - 1));
      dst16[pos] = static_cast<uint16_t>(v & (kSpan - 1));
    }
  }
}

}  // namespace

void build_layout(gx_graph* g) {
  DegreeOrder& D = g->degree_order;
  PbLayout& L = D.pb;
  if (L.ready) return;
  const int64_t n = g->n, m = g->m;
  GX_REQUIRE(D.ready, "propagation blocking needs the degree order");
  GX_REQUIRE(supported(n, m), "graph too large or empty for the propagation-blocked sweep");
  hipStream_t s = g->stream;
  const int64_t nb = (n + kSpan - 1) / kSpan;
  const int64_t nkeys = nb * nb;

  std::vector<uint32_t> start(nkeys + 1);
  {
    DevBuf esrc, cnt, cursor;
    esrc.ensure(m * 4);
    edge_sources(g->row, n, m, esrc.as<uint32_t>(), s);
    const uint32_t* rank = D.new_of_old.as<uint32_t>();
    cnt.ensure((nkeys + 1) * 4);
    cursor.ensure((nkeys + 1) * 4);
    GX_HIP(hipMemsetAsync(cnt.p, 0, (nkeys + 1) * 4, s));
    const unsigned grid = grid_for(m, kBuildBlock, 1 << 16);
    count_kernel<<<grid, kBuildBlock, 0, s>>>(esrc.as<uint32_t>(), g->col, rank, m, static_cast<uint32_t>(nb),
                                               cnt.as<uint32_t>());
    GX_LAUNCHED("pb count_kernel");
    exclusive_scan_u32(cnt.as<uint32_t>(), cursor.as<uint32_t>(), nkeys + 1, s);
    GX_HIP(hipStreamSynchronize(s));
    GX_HIP(hipMemcpy(start.data(), cursor.p, (nkeys + 1) * 4, hipMemcpyDeviceToHost));
    GX_REQUIRE(start[nkeys] == static_cast<uint32_t>(m), "propagation-blocking layout: edge count mismatch");
    L.src16.ensure(m * 2);
    L.dst16.ensure(m * 2);
    scatter_kernel<<<grid, kBuildBlock, 0, s>>>(esrc.as<uint32_t>(), g->col, rank, m, static_cast<uint32_t>(nb),
                                                 cursor.as<uint32_t>(),
                                                 L.src16.as<uint16_t>(), L.dst16.as<uint16_t>());
    GX_LAUNCHED("pb scatter_kernel");
    GX_HIP(hipStreamSynchronize(s));
  }

  // Phase-1 pieces: every nonempty chunk in slices of kPieceEntries.
  std::vector<Piece> pieces;
  for (int64_t c = 0; c < nb; ++c) {
    const uint32_t a = start[c * nb], b = start[(c + 1) * nb];
    for (uint64_t p = a; p < b; p += kPieceEntries)
      pieces.push_back(Piece{static_cast<uint32_t>(c), static_cast<uint32_t>(p),
                             static_cast<uint32_t>(std::min<uint64_t>(b, p + kPieceEntries)), 0u});
  }

  // Phase-2 jobs per bin and class (chunk-major scan, appended per bin).
  std::vector<std::vector<Job>> per[3];
  for (auto& v : per) v.resize(nb);
  for (int64_t c = 0; c < nb; ++c) {
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t key = c * nb + b;
      uint32_t p = start[key];
      uint32_t len = start[key + 1] - p;
      while (len > 0) {
        const uint32_t l = std::min<uint32_t>(len, kJob);
        const int cls = l > kMid ? 0 : (l > kShort ? 1 : 2);
        per[cls][b].push_back(Job{p, l});
        p += l;
        len -= l;
      }
    }
  }
  std::vector<Job> jobs[3];
  std::vector<uint32_t> first[3];  // first job of every bin in each class list
  for (int cls = 0; cls < 3; ++cls) {
    first[cls].resize(nb + 1);
    for (int64_t b = 0; b < nb; ++b) {
      first[cls][b] = static_cast<uint32_t>(jobs[cls].size());
      jobs[cls].insert(jobs[cls].end(), per[cls][b].begin(), per[cls][b].end());
      std::vector<Job>().swap(per[cls][b]);
    }
    first[cls][nb] = static_cast<uint32_t>(jobs[cls].size());
  }

  // Tasks: a bin with at most kTaskEntries edges is one task; a heavier bin
  // is cut into slices of about kTaskEntries (long, then mid, then short
  // jobs) that share a global accumulator slot.
  std::vector<std::pair<uint64_t, Task>> tasks;
  std::vector<Multi> multi;
  for (int64_t b = 0; b < nb; ++b) {
    uint64_t entries = 0;
    for (int cls = 0; cls < 3; ++cls)
      for (uint32_t j = first[cls][b]; j < first[cls][b + 1]; ++j) entries += jobs[cls][j].len;
    Task t{};
    t.bin = static_cast<uint32_t>(b);
    if (entries <= static_cast<uint64_t>(kTaskEntries)) {
      t.slot = kSingle;
      for (int cls = 0; cls < 3; ++cls) {
        t.job0[cls] = first[cls][b];
        t.job1[cls] = first[cls][b + 1];
      }
      tasks.push_back({entries, t});
      continue;
    }
    t.slot = static_cast<uint32_t>(multi.size());
    multi.push_back(Multi{static_cast<uint32_t>(b), t.slot});
    // A slice starts at job j of class cls: earlier classes are done, later
    // ones start at the bin's first job; every range is empty until filled.
    auto reset = [&](int cls, uint32_t j) {
      for (int k = 0; k < 3; ++k)
        t.job0[k] = t.job1[k] = k < cls ? first[k][b + 1] : (k == cls ? j : first[k][b]);
    };
    reset(0, first[0][b]);
    uint64_t acc = 0;
    for (int cls = 0; cls < 3; ++cls) {
      for (uint32_t j = first[cls][b]; j < first[cls][b + 1]; ++j) {
        t.job1[cls] = j + 1;
        acc += jobs[cls][j].len;
        if (acc >= static_cast<uint64_t>(kTaskEntries)) {
          tasks.push_back({acc, t});
          acc = 0;
          reset(cls, j + 1);
        }
      }
    }
    if (acc > 0) tasks.push_back({acc, t});
  }
  std::stable_sort(tasks.begin(), tasks.end(),
                   [](const std::pair<uint64_t, Task>& a, const std::pair<uint64_t, Task>& b) {
                     return a.first > b.first;
                   });

  // The tables must cover every edge exactly once, every job range must lie
  // in its list and every bin must be finalised once, before any launch.
  {
    uint64_t covered = 0, pieced = 0;
    std::vector<uint8_t> fin(nb, 0);
    for (const auto& pt : tasks) {
      const Task& t = pt.second;
      uint64_t e = 0;
      for (int cls = 0; cls < 3; ++cls) {
        GX_REQUIRE(t.job0[cls] <= t.job1[cls] && t.job1[cls] <= jobs[cls].size() &&
                       t.job0[cls] >= first[cls][t.bin] && t.job1[cls] <= first[cls][t.bin + 1],
                   "propagation-blocking layout: bad job range");
        for (uint32_t j = t.job0[cls]; j < t.job1[cls]; ++j) e += jobs[cls][j].len;
      }
      GX_REQUIRE(e == pt.first, "propagation-blocking layout: task size mismatch");
      covered += e;
      if (t.slot == kSingle) fin[t.bin] += 1;
      else GX_REQUIRE(t.slot < multi.size() && multi[t.slot].bin == t.bin, "propagation-blocking layout: bad slot");
    }
    for (const Multi& mu : multi) fin[mu.bin] += 1;
    for (int64_t b = 0; b < nb; ++b) GX_REQUIRE(fin[b] == 1, "propagation-blocking layout: bin not finalised once");
    for (const Piece& pc : pieces) {
      GX_REQUIRE(pc.p0 < pc.p1 && pc.p1 <= static_cast<uint64_t>(m) && pc.chunk < nb,
                 "propagation-blocking layout: bad piece");
      pieced += pc.p1 - pc.p0;
    }
    GX_REQUIRE(covered == static_cast<uint64_t>(m) && pieced == static_cast<uint64_t>(m),
               "propagation-blocking layout: edges not covered exactly once");
  }

  // Upload.
  std::vector<Task> tv(tasks.size());
  for (size_t i = 0; i < tasks.size(); ++i) tv[i] = tasks[i].second;
  L.pieces.ensure(std::max<size_t>(pieces.size(), 1) * sizeof(Piece));
  L.tasks.ensure(tv.size() * sizeof(Task));
  L.multi.ensure(std::max<size_t>(multi.size(), 1) * sizeof(Multi));
  const size_t nj = jobs[0].size() + jobs[1].size() + jobs[2].size();
  L.jobs.ensure(std::max<size_t>(nj, 1) * sizeof(Job));
  GX_HIP(hipMemcpy(L.pieces.p, pieces.data(), pieces.size() * sizeof(Piece), hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(L.tasks.p, tv.data(), tv.size() * sizeof(Task), hipMemcpyHostToDevice));
  if (!multi.empty())
    GX_HIP(hipMemcpy(L.multi.p, multi.data(), multi.size() * sizeof(Multi), hipMemcpyHostToDevice));
  size_t off = 0;
  for (int cls = 0; cls < 3; ++cls) {
    if (!jobs[cls].empty())
      GX_HIP(hipMemcpy(L.jobs.as<Job>() + off, jobs[cls].data(), jobs[cls].size() * sizeof(Job),
                       hipMemcpyHostToDevice));
    L.njobs[cls] = static_cast<int64_t>(jobs[cls].size());
    off += jobs[cls].size();
  }
  L.gacc.ensure(std::max<size_t>(multi.size(), 1) * kSpan * sizeof(unsigned long long));
  GX_HIP(hipMemsetAsync(L.gacc.p, 0, std::max<size_t>(multi.size(), 1) * kSpan * sizeof(unsigned long long), s));
  L.vals.ensure(m * sizeof(float));
  GX_HIP(hipStreamSynchronize(s));
  L.nbins = nb;
  L.npieces = static_cast<int64_t>(pieces.size());
  L.ntasks = static_cast<int64_t>(tv.size());
  L.nmulti = static_cast<int64_t>(multi.size());
  L.ready = true;
}

}  // namespace pb
}  // namespace gx
