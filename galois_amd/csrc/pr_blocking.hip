// Layout and work tables of the propagation-blocked PageRank sweep
// (pr_blocking.hpp), built once per graph from the CSR and the degree order.
//
//  1. key_kernel: one 64-bit key per edge, (run, source in chunk, destination
//     in bin) in degree-order ids, run = chunk of the source * bins + bin of
//     the destination; a radix sort over the used bits puts the edges in
//     layout order (deterministic: the graph has no duplicate edges).
//  2. mark_kernel: dst16 with the group flag, the flag array for the group
//     numbering, and the edge and group counts per run (one atomic per run of
//     equal keys in a wavefront).
//  3. exclusive scans: group number of every edge's group start, start of
//     every run in edges and in groups; bin_major_kernel moves dst16 into
//     bin-major run order (the runs of one bin, from every chunk, in one
//     stretch, long runs first, then mid, then short: each of phase 2's
//     three passes over a bin reads one contiguous part); group_kernel
//     writes gsrc16 (groups stay in run-major order, as phase 1 writes them).
//  4. host: phase-1 pieces per chunk (in groups), phase-2 jobs per bin in
//     three length classes (start in bin-major dst16), tasks (bins, heavy
//     bins sliced), sorted longest first; job_base_kernel then fills every
//     job's group base from its run-major first edge.
//
// Edges within a run are in (source, destination) order: neighbouring edges
// of a group have distinct, ascending destinations, so a wavefront's adds
// mostly hit different LDS words.
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

__global__ void __launch_bounds__(kBuildBlock) key_kernel(const uint32_t* __restrict__ esrc,
                                                          const int32_t* __restrict__ col,
                                                          const uint32_t* __restrict__ rank, int64_t m, uint32_t nb,
                                                          uint64_t* __restrict__ keys) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(kBuildBlock) + threadIdx.x; e < m; e += stride) {
    const uint32_t u = rank[esrc[e]], v = rank[col[e]];
    const uint64_t run = static_cast<uint64_t>(u >> kShift) * nb + (v >> kShift);
    keys[e] = (run << (2 * kShift)) | (static_cast<uint64_t>(u & (kSpan - 1)) << kShift) | (v & (kSpan - 1));
  }
}

// Sorted keys -> dst16 (+ group flag), flag[e] (0/1), edges and groups per run.
__global__ void __launch_bounds__(kBuildBlock) mark_kernel(const uint64_t* __restrict__ keys, int64_t m,
                                                           uint16_t* __restrict__ dst16, uint32_t* __restrict__ flag,
                                                           uint32_t* __restrict__ cnt, uint32_t* __restrict__ gcnt) {
  const unsigned lane = lane_id();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e0 = blockIdx.x * static_cast<int64_t>(kBuildBlock) + (threadIdx.x & ~(kWave - 1)); e0 < m;
       e0 += stride) {
    const int64_t e = e0 + lane;
    const bool valid = e < m;
    const unsigned nvalid = static_cast<unsigned>(m - e0 < kWave ? m - e0 : kWave);
    const uint64_t k = valid ? keys[e] : 0ull;
    const bool first = valid && (e == 0 || (keys[e - 1] >> kShift) != (k >> kShift));
    const uint32_t run = valid ? static_cast<uint32_t>(k >> (2 * kShift)) : 0xffffffffu;
    if (valid) {
      dst16[e] = static_cast<uint16_t>((k & (kSpan - 1)) | (first ? kFirst : 0u));
      flag[e] = first ? 1u : 0u;
    }
    unsigned len;
    const uint64_t heads = run_heads(valid, run, lane, nvalid, len);
    const uint64_t f = __ballot(first);
    if ((heads >> lane) & 1ull) {
      const uint64_t span = (len >= 64 ? ~0ull : ((uint64_t(1) << len) - 1)) << lane;
      atomicAdd(&cnt[run], len);
      atomicAdd(&gcnt[run], static_cast<uint32_t>(__popcll(f & span)));
    }
  }
}

// gsrc16[group] = source in chunk, for every edge that opens a group.
__global__ void __launch_bounds__(kBuildBlock) group_kernel(const uint64_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ flag,
                                                            const uint32_t* __restrict__ gpos, int64_t m,
                                                            uint16_t* __restrict__ gsrc16) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(kBuildBlock) + threadIdx.x; e < m; e += stride)
    if (flag[e]) gsrc16[gpos[e]] = static_cast<uint16_t>((keys[e] >> kShift) & (kSpan - 1));
}

// dst16 from run-major (chunk, bin) order into bin-major (bin, chunk)
// order: edge e of run r moves from cstart[r] + i to bstart[r] + i.
__global__ void __launch_bounds__(kBuildBlock) bin_major_kernel(const uint64_t* __restrict__ keys, int64_t m,
                                                                const uint32_t* __restrict__ cstart,
                                                                const uint32_t* __restrict__ bstart,
                                                                const uint16_t* __restrict__ in,
                                                                uint16_t* __restrict__ out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t e = blockIdx.x * static_cast<int64_t>(kBuildBlock) + threadIdx.x; e < m; e += stride) {
    const uint32_t r = static_cast<uint32_t>(keys[e] >> (2 * kShift));
    out[bstart[r] + (static_cast<uint32_t>(e) - cstart[r])] = in[e];
  }
}

// gbase of every job: groups opened before its first edge, minus one.  The
// gbase field holds the job's first edge in run-major order on entry.
__global__ void __launch_bounds__(kBuildBlock) job_base_kernel(Job* __restrict__ jobs, int64_t njobs,
                                                               const uint32_t* __restrict__ gpos) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBuildBlock;
  for (int64_t j = blockIdx.x * static_cast<int64_t>(kBuildBlock) + threadIdx.x; j < njobs; j += stride)
    jobs[j].gbase = gpos[jobs[j].gbase] - 1u;
}

// Number of vertices with out-edges: out-degrees are non-increasing in
// degree order, so it is the first index whose out-degree is 0 (one thread,
// binary search).
__global__ void nonzero_kernel(const uint32_t* __restrict__ outdeg, int64_t n, int64_t* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (outdeg[mid] > 0) lo = mid + 1; else hi = mid;
  }
  *out = lo;
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

  std::vector<uint32_t> start(nkeys + 1), gstart(nkeys + 1), bstart;
  DevBuf gpos;  // exclusive scan of the group flags, m + 1 entries
  int64_t ngroups = 0;
  {
    DevBuf keys, sorted;
    keys.ensure(m * 8);
    {
      DevBuf esrc;
      esrc.ensure(m * 4);
      edge_sources(g->row, n, m, esrc.as<uint32_t>(), s);
      const uint32_t* rank = D.new_of_old.as<uint32_t>();
      key_kernel<<<grid_for(m, kBuildBlock, 1 << 16), kBuildBlock, 0, s>>>(
          esrc.as<uint32_t>(), g->col, rank, m, static_cast<uint32_t>(nb), keys.as<uint64_t>());
      GX_LAUNCHED("pb key_kernel");
    }
    int run_bits = 0;
    while ((int64_t(1) << run_bits) < nkeys) ++run_bits;
    sorted.ensure(m * 8);
    sort_keys_u64(keys.as<uint64_t>(), sorted.as<uint64_t>(), m, 2 * kShift + run_bits, s);
    keys.release();
    DevBuf flag, cnt, gcnt, cur, gcur;
    flag.ensure((m + 1) * 4);
    cnt.ensure((nkeys + 1) * 4);
    gcnt.ensure((nkeys + 1) * 4);
    cur.ensure((nkeys + 1) * 4);
    gcur.ensure((nkeys + 1) * 4);
    gpos.ensure((m + 1) * 4);
    DevBuf d16;  // run-major dst16; L.dst16 gets it bin-major below
    d16.ensure(m * 2);
    GX_HIP(hipMemsetAsync(cnt.p, 0, (nkeys + 1) * 4, s));
    GX_HIP(hipMemsetAsync(gcnt.p, 0, (nkeys + 1) * 4, s));
    GX_HIP(hipMemsetAsync(flag.as<uint32_t>() + m, 0, 4, s));
    mark_kernel<<<grid_for(m, kBuildBlock, 1 << 16), kBuildBlock, 0, s>>>(
        sorted.as<uint64_t>(), m, d16.as<uint16_t>(), flag.as<uint32_t>(), cnt.as<uint32_t>(),
        gcnt.as<uint32_t>());
    GX_LAUNCHED("pb mark_kernel");
    exclusive_scan_u32(cnt.as<uint32_t>(), cur.as<uint32_t>(), nkeys + 1, s);
    exclusive_scan_u32(gcnt.as<uint32_t>(), gcur.as<uint32_t>(), nkeys + 1, s);
    exclusive_scan_u32(flag.as<uint32_t>(), gpos.as<uint32_t>(), m + 1, s);
    GX_HIP(hipStreamSynchronize(s));
    GX_HIP(hipMemcpy(start.data(), cur.p, (nkeys + 1) * 4, hipMemcpyDeviceToHost));
    GX_HIP(hipMemcpy(gstart.data(), gcur.p, (nkeys + 1) * 4, hipMemcpyDeviceToHost));
    uint32_t total = 0;
    GX_HIP(hipMemcpy(&total, gpos.as<uint32_t>() + m, 4, hipMemcpyDeviceToHost));
    ngroups = total;
    GX_REQUIRE(start[nkeys] == static_cast<uint32_t>(m), "propagation-blocking layout: edge count mismatch");
    GX_REQUIRE(gstart[nkeys] == total && ngroups >= 1 && ngroups <= m,
               "propagation-blocking layout: group count mismatch");
    // Bin-major run offsets: phase 2 reads the runs of one bin, from every
    // chunk, as one stretch of dst16, and inside it the runs whose first job
    // is long, then mid, then short, so that each of its three passes over
    // the bin (one per job class) reads a contiguous part.
    // Two passes in run order (the table is 4 bytes x bins^2): edges per
    // (bin, class), their exclusive scan, then a cursor per (bin, class) that
    // the chunks, in ascending order, advance.
    bstart.resize(nkeys + 1);
    {
      auto cls_of = [](uint32_t len) {
        const uint32_t l = std::min<uint32_t>(len, kJob);
        return l > kMid ? 0 : (l > kShort ? 1 : 2);
      };
      std::vector<uint32_t> cursor(nb * 3, 0);
      for (int64_t key = 0; key < nkeys; ++key) {
        const uint32_t len = start[key + 1] - start[key];
        cursor[(key % nb) * 3 + cls_of(len)] += len;
      }
      uint32_t at = 0;
      for (uint32_t& v : cursor) {
        const uint32_t t = v;
        v = at;
        at += t;
      }
      GX_REQUIRE(at == static_cast<uint32_t>(m), "propagation-blocking layout: bin-major offsets");
      for (int64_t key = 0; key < nkeys; ++key) {
        const uint32_t len = start[key + 1] - start[key];
        uint32_t& p = cursor[(key % nb) * 3 + cls_of(len)];
        bstart[key] = p;
        p += len;
      }
      bstart[nkeys] = at;
    }
    {
      DevBuf bst;
      bst.ensure(nkeys * 4);
      GX_HIP(hipMemcpy(bst.p, bstart.data(), nkeys * 4, hipMemcpyHostToDevice));
      L.dst16.ensure(m * 2);
      bin_major_kernel<<<grid_for(m, kBuildBlock, 1 << 16), kBuildBlock, 0, s>>>(
          sorted.as<uint64_t>(), m, cur.as<uint32_t>(), bst.as<uint32_t>(), d16.as<uint16_t>(),
          L.dst16.as<uint16_t>());
      GX_LAUNCHED("pb bin_major_kernel");
      GX_HIP(hipStreamSynchronize(s));
    }
    L.gsrc16.ensure(ngroups * 2);
    group_kernel<<<grid_for(m, kBuildBlock, 1 << 16), kBuildBlock, 0, s>>>(
        sorted.as<uint64_t>(), flag.as<uint32_t>(), gpos.as<uint32_t>(), m, L.gsrc16.as<uint16_t>());
    GX_LAUNCHED("pb group_kernel");
    GX_HIP(hipStreamSynchronize(s));
  }

  // Phase-1 pieces: the groups of every nonempty chunk in slices of
  // kPieceEntries.
  std::vector<Piece> pieces;
  for (int64_t c = 0; c < nb; ++c) {
    const uint32_t a = gstart[c * nb], b = gstart[(c + 1) * nb];
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
        per[cls][b].push_back(Job{bstart[key] + (p - start[key]), l, p});  // gbase: set on the device
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

  // Tasks: a bin with at most `task_entries` edges (knob "pr.task_edges")
  // is one task; a heavier bin is cut into slices of about that many (long,
  // then mid, then short jobs) that share a global accumulator slot.  The
  // default, m / 1024 within [2^19, 2^20] (pr_task_edges_default), grows the
  // tasks of graphs with more than 2^29 edges only, which still get about
  // four tasks per CU.  A slice closes at the first job that reaches the
  // limit, so it holds up to kJob - 1 edges more.
  const uint64_t task_entries = static_cast<uint64_t>(tune_effective("pr.task_edges", n, m));
  std::vector<std::pair<uint64_t, Task>> tasks;
  std::vector<Multi> multi;
  for (int64_t b = 0; b < nb; ++b) {
    uint64_t entries = 0;
    for (int cls = 0; cls < 3; ++cls)
      for (uint32_t j = first[cls][b]; j < first[cls][b + 1]; ++j) entries += jobs[cls][j].len;
    Task t{};
    t.bin = static_cast<uint32_t>(b);
    if (entries <= task_entries) {
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
        if (acc >= task_entries) {
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
      GX_REQUIRE(pc.p0 < pc.p1 && pc.p1 <= static_cast<uint64_t>(ngroups) && pc.chunk < nb,
                 "propagation-blocking layout: bad piece");
      pieced += pc.p1 - pc.p0;
    }
    GX_REQUIRE(covered == static_cast<uint64_t>(m) && pieced == static_cast<uint64_t>(ngroups),
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
  if (nj > 0) {
    job_base_kernel<<<grid_for(static_cast<int64_t>(nj), kBuildBlock, 1 << 14), kBuildBlock, 0, s>>>(
        L.jobs.as<Job>(), static_cast<int64_t>(nj), gpos.as<uint32_t>());
    GX_LAUNCHED("pb job_base_kernel");
  }
  L.gacc.ensure(std::max<size_t>(multi.size(), 1) * kSpan * sizeof(unsigned long long));
  GX_HIP(hipMemsetAsync(L.gacc.p, 0, std::max<size_t>(multi.size(), 1) * kSpan * sizeof(unsigned long long), s));
  L.vals.ensure(ngroups * sizeof(float));
  GX_HIP(hipStreamSynchronize(s));
  L.ngroups = ngroups;
  L.nbins = nb;
  L.npieces = static_cast<int64_t>(pieces.size());
  L.ntasks = static_cast<int64_t>(tv.size());
  L.nmulti = static_cast<int64_t>(multi.size());
  L.task_entries = static_cast<int64_t>(task_entries);
  {
    DevBuf nz;
    nz.ensure(8);
    nonzero_kernel<<<1, 64, 0, s>>>(D.outdeg.as<uint32_t>(), n, nz.as<int64_t>());
    GX_LAUNCHED("pb nonzero_kernel");
    GX_HIP(hipMemcpy(&L.nonzero, nz.p, 8, hipMemcpyDeviceToHost));
  }
  L.ready = true;
}

}  // namespace pb
}  // namespace gx
