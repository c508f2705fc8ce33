// ORACLE — test infrastructure only.
//
// A CPU restatement of the Galois lonestar analytics algorithms (the
// reference's own CPU path) used (1) by tests/ as the parity checker of the
// HIP library and (2) by bench.py's cpu_baseline leg as the timed CPU
// baseline.  Nothing in the product (galois_amd/) links, imports or calls
// this code.  It is built into oracle/liboracle.so by oracle/Makefile.
//
// Each function cites the reference lines it follows (paths relative to the
// Galois source tree).  The reference itself cannot be compiled here: it
// needs Boost, LLVM's CommandLine library and fmt, none of which exist in
// this image.  Parity of each function is pinned in tests/test_oracle.py
// against golden vectors written by networkx/scipy/brute force
// (tests/golden/) and against closed forms and invariants; see the function
// comments for which check pins which function.
//
// Parallel loops use a small persistent std::thread pool with dynamic
// chunked scheduling, the role of galois::do_all(steal(), chunk_size<K>).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kInf = 2147483646u;  // BFS_SSSP.h:31-32: UINT32_MAX/2 - 1

// ------------------------------------------------------------ thread pool
class Pool {
 public:
  explicit Pool(int threads) : nthreads_(std::max(1, threads)) {
    for (int t = 1; t < nthreads_; ++t) workers_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int size() const { return nthreads_; }
  // Run fn(tid) on every thread (tid 0 is the caller) and wait.
  void on_each(const std::function<void(int)>& fn) {
    if (nthreads_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      pending_ = nthreads_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  void loop(int tid) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        job = job_;
      }
      (*job)(tid);
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (--pending_ == 0) done_cv_.notify_one();
      }
    }
  }
  int nthreads_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// galois::do_all(iterate(begin,end), f, steal(), chunk_size<chunk>())
template <class F>
void do_all(Pool& pool, int64_t begin, int64_t end, int64_t chunk, F&& f) {
  if (end <= begin) return;
  if (pool.size() == 1 || end - begin <= chunk) {
    for (int64_t i = begin; i < end; ++i) f(i, 0);
    return;
  }
  std::atomic<int64_t> next{begin};
  pool.on_each([&](int tid) {
    for (;;) {
      int64_t s = next.fetch_add(chunk, std::memory_order_relaxed);
      if (s >= end) break;
      int64_t e = std::min(s + chunk, end);
      for (int64_t i = s; i < e; ++i) f(i, tid);
    }
  });
}

inline uint32_t atomic_min_u32(std::atomic<uint32_t>& a, uint32_t v) {
  // galois::atomicMin (libgalois/include/galois/AtomicHelpers.h:60-67):
  // returns the old value.
  uint32_t old = a.load(std::memory_order_relaxed);
  while (old > v && !a.compare_exchange_weak(old, v, std::memory_order_relaxed)) {
  }
  return old;
}

// In-edges (CSC) of a CSR with sources ascending inside each column: the
// tools/graph-convert transpose (graph-convert.cpp:1831-1880).
void transpose(int64_t n, const int32_t* row, const int32_t* col, std::vector<int32_t>& trow,
               std::vector<int32_t>& tcol, std::vector<int32_t>* perm) {
  const int64_t m = row[n];
  trow.assign(n + 1, 0);
  tcol.assign(m, 0);
  if (perm) perm->assign(m, 0);
  for (int64_t e = 0; e < m; ++e) trow[col[e] + 1]++;
  for (int64_t v = 0; v < n; ++v) trow[v + 1] += trow[v];
  std::vector<int32_t> cur(trow.begin(), trow.end() - 1);
  for (int64_t u = 0; u < n; ++u)
    for (int32_t e = row[u]; e < row[u + 1]; ++e) {
      int32_t slot = cur[col[e]]++;
      tcol[slot] = static_cast<int32_t>(u);
      if (perm) (*perm)[slot] = e;
    }
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ BFS
// Level-synchronous BFS, lonestar/analytics/cpu/bfs/bfs.cpp:245-293
// (syncAlgo): level[src] = 0; per level a do_all over the frontier with
// chunk 256 (bfs.cpp:85); dst gets level L+1 when still DIST_INFINITY.  The
// parallel write race of the reference (same value) is resolved with a CAS so
// each vertex enters the next frontier once.  Pinned by networkx
// single_source_shortest_path_length fixtures and the BFS_SSSP::verify
// invariants (BFS_SSSP.h:231-263) in tests/test_oracle.py.
int oracle_bfs(int64_t n, const int32_t* row, const int32_t* col, uint32_t src, uint32_t* level_out,
               int threads) {
  if (src >= n) return 1;
  Pool pool(threads);
  std::vector<std::atomic<uint32_t>> level(n);
  for (int64_t v = 0; v < n; ++v) level[v].store(kInf, std::memory_order_relaxed);
  level[src] = 0;
  std::vector<uint32_t> cur{src};
  std::vector<std::vector<uint32_t>> local(pool.size());
  uint32_t L = 0;
  while (!cur.empty()) {
    ++L;
    for (auto& l : local) l.clear();
    do_all(pool, 0, static_cast<int64_t>(cur.size()), 256, [&](int64_t i, int tid) {
      uint32_t u = cur[i];
      for (int32_t e = row[u]; e < row[u + 1]; ++e) {
        uint32_t d = static_cast<uint32_t>(col[e]);
        if (level[d].load(std::memory_order_relaxed) == kInf) {
          uint32_t expect = kInf;
          if (level[d].compare_exchange_strong(expect, L, std::memory_order_relaxed)) local[tid].push_back(d);
        }
      }
    });
    cur.clear();
    for (auto& l : local) cur.insert(cur.end(), l.begin(), l.end());
  }
  for (int64_t v = 0; v < n; ++v) level_out[v] = level[v].load(std::memory_order_relaxed);
  return 0;
}

// ------------------------------------------------------------------ SSSP
// Delta-stepping, lonestar/analytics/cpu/sssp/SSSP.cpp:119-173 with the
// OBIM bucket index dist >> shift (BFS_SSSP.h:50-58): buckets are drained in
// increasing order; a popped item whose vertex already has a smaller
// distance is dropped (SSSP.cpp:138); relaxations use atomicMin and push
// {dst, newDist} when they lower the distance.  Within a bucket, items are
// processed in parallel rounds (the barrier variant, SSSP.cpp OBIM_Barrier).
// Shortest distances are unique, so the result is schedule independent.
// Pinned by scipy.sparse.csgraph.dijkstra fixtures and by oracle_dijkstra.
int oracle_sssp_delta(int64_t n, const int32_t* row, const int32_t* col, const uint32_t* w, uint32_t src,
                      uint32_t shift, uint32_t* dist_out, int threads) {
  if (src >= n) return 1;
  Pool pool(threads);
  std::vector<std::atomic<uint32_t>> dist(n);
  for (int64_t v = 0; v < n; ++v) dist[v].store(kInf, std::memory_order_relaxed);
  dist[src] = 0;
  struct Item {
    uint32_t v, d;
  };
  std::vector<std::vector<Item>> buckets(1);
  buckets[0].push_back({src, 0});
  struct Local {
    std::vector<Item> same;
    std::vector<Item> later;
  };
  std::vector<Local> local(pool.size());
  for (size_t b = 0; b < buckets.size(); ++b) {
    std::vector<Item> items;
    items.swap(buckets[b]);
    while (!items.empty()) {
      for (auto& l : local) {
        l.same.clear();
        l.later.clear();
      }
      do_all(pool, 0, static_cast<int64_t>(items.size()), 64, [&](int64_t i, int tid) {
        const Item it = items[i];
        const uint32_t sd = dist[it.v].load(std::memory_order_relaxed);
        if (sd < it.d) return;  // stale work item
        for (int32_t e = row[it.v]; e < row[it.v + 1]; ++e) {
          uint32_t d = static_cast<uint32_t>(col[e]);
          uint32_t nd = sd + w[e];
          uint32_t old = atomic_min_u32(dist[d], nd);
          if (nd < old) {
            if ((nd >> shift) == b)
              local[tid].same.push_back({d, nd});
            else
              local[tid].later.push_back({d, nd});
          }
        }
      });
      items.clear();
      for (auto& l : local) {
        items.insert(items.end(), l.same.begin(), l.same.end());
        for (const Item& it : l.later) {
          size_t nb = it.d >> shift;
          if (nb >= buckets.size()) buckets.resize(nb + 1);
          buckets[nb].push_back(it);
        }
      }
    }
  }
  for (int64_t v = 0; v < n; ++v) dist_out[v] = dist[v].load(std::memory_order_relaxed);
  return 0;
}

// Serial Dijkstra, SSSP.cpp:224-261 (min-heap of {dist, node}, stale pops
// skipped).  Pinned by scipy dijkstra fixtures.
int oracle_dijkstra(int64_t n, const int32_t* row, const int32_t* col, const uint32_t* w, uint32_t src,
                    uint32_t* dist) {
  if (src >= n) return 1;
  for (int64_t v = 0; v < n; ++v) dist[v] = kInf;
  dist[src] = 0;
  using P = std::pair<uint32_t, uint32_t>;
  std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
  pq.push({0, src});
  while (!pq.empty()) {
    auto [d, u] = pq.top();
    pq.pop();
    if (dist[u] < d) continue;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) {
      uint32_t v = static_cast<uint32_t>(col[e]);
      uint32_t nd = d + w[e];
      if (nd < dist[v]) {
        dist[v] = nd;
        pq.push({nd, v});
      }
    }
  }
  return 0;
}

// ------------------------------------------------------------------- CC
// Union-find connected components: lonestar/analytics/cpu/
// connected-components/ConnectedComponents.cpp:341-383 (union over edges
// with src < dst, then compress) with the lock-free UnionFindNode of
// libgalois/include/galois/UnionFind.h:32-123 (merge links the higher root
// into the lower one, :108-122).  After compression every vertex holds the
// minimum vertex id of its component.  Pinned by
// scipy.sparse.csgraph.connected_components fixtures (relabelled to the
// component minimum) and the verify invariant (ConnectedComponents.cpp:
// 1046-1072).
int oracle_cc(int64_t n, const int32_t* row, const int32_t* col, uint32_t* label, int threads) {
  Pool pool(threads);
  std::vector<std::atomic<uint32_t>> parent(n);
  for (int64_t v = 0; v < n; ++v) parent[v].store(static_cast<uint32_t>(v), std::memory_order_relaxed);
  auto find = [&](uint32_t x) {
    for (;;) {
      uint32_t p = parent[x].load(std::memory_order_relaxed);
      if (p == x) return x;
      uint32_t gp = parent[p].load(std::memory_order_relaxed);
      if (gp != p) parent[x].compare_exchange_weak(p, gp, std::memory_order_relaxed);  // path halving
      x = gp;
    }
  };
  do_all(pool, 0, n, 512, [&](int64_t u, int) {
    for (int32_t e = row[u]; e < row[u + 1]; ++e) {
      uint32_t v = static_cast<uint32_t>(col[e]);
      if (static_cast<int64_t>(v) <= u) continue;  // each undirected edge once
      uint32_t a = static_cast<uint32_t>(u), b = v;
      for (;;) {
        a = find(a);
        b = find(b);
        if (a == b) break;
        if (a < b) std::swap(a, b);  // a = higher root
        uint32_t expect = a;
        if (parent[a].compare_exchange_strong(expect, b, std::memory_order_relaxed)) break;
      }
    }
  });
  do_all(pool, 0, n, 512, [&](int64_t v, int) { label[v] = find(static_cast<uint32_t>(v)); });
  return 0;
}

// ---------------------------------------------------------------- k-core
// lonestar/analytics/cpu/k-core/kcore.cpp:86-95 (degreeCounting), :103-115
// (initial dead worklist: degree < k), :124-159 (syncCascadeKCore: each dead
// vertex decrements every neighbour with atomicSubtract and a neighbour whose
// old degree was exactly k joins the next worklist), :201-217 (alive =
// currentDegree >= k).  The final degree array is the number of alive
// neighbours and is schedule independent.  Pinned by networkx.k_core
// fixtures and the serial brute-force peel in tests.
int oracle_kcore(int64_t n, const int32_t* row, const int32_t* col, uint32_t k, uint32_t* degree_out,
                 int threads) {
  Pool pool(threads);
  std::vector<std::atomic<uint32_t>> deg(n);
  for (int64_t v = 0; v < n; ++v) deg[v].store(static_cast<uint32_t>(row[v + 1] - row[v]), std::memory_order_relaxed);
  std::vector<std::vector<uint32_t>> local(pool.size());
  std::vector<uint32_t> cur;
  for (int64_t v = 0; v < n; ++v)
    if (deg[v].load(std::memory_order_relaxed) < k) cur.push_back(static_cast<uint32_t>(v));
  while (!cur.empty()) {
    for (auto& l : local) l.clear();
    do_all(pool, 0, static_cast<int64_t>(cur.size()), 64, [&](int64_t i, int tid) {
      uint32_t dead = cur[i];
      for (int32_t e = row[dead]; e < row[dead + 1]; ++e) {
        uint32_t d = static_cast<uint32_t>(col[e]);
        uint32_t old = deg[d].fetch_sub(1u, std::memory_order_relaxed);
        if (old == k) local[tid].push_back(d);
      }
    });
    cur.clear();
    for (auto& l : local) cur.insert(cur.end(), l.begin(), l.end());
  }
  for (int64_t v = 0; v < n; ++v) degree_out[v] = deg[v].load(std::memory_order_relaxed);
  return 0;
}

// -------------------------------------------------------------- PageRank
// Normalised PageRank with uniform teleport and dangling mass redistributed
// uniformly, FP64 Jacobi power iteration pulling over the transpose (the
// loop structure of computePRTopological, PageRank-pull.cpp:196-251, without
// its in-place update):
//   x0 = 1/n;  x'[v] = (1-a)/n + a*(sum_{u->v} x[u]/outdeg(u) + D/n),
//   D = sum_{outdeg(u)=0} x[u];  stop when sum|x'-x| < tol or max_iter.
// Pinned against a direct sparse solve of (I - a*M^T) x = (1-a)/n (scipy
// spsolve fixtures) and sum(x) = 1.
int oracle_pagerank_fp64_csc(int64_t n, const uint32_t* outdeg, const int32_t* trow, const int32_t* tcol,
                             double alpha, double tol, uint32_t max_iter, double* x_out, uint32_t* iters_out,
                             double* resid_out, int threads) {
  Pool pool(threads);
  std::vector<double> x(n, 1.0 / static_cast<double>(n)), y(n), xn(n);
  uint32_t it = 0;
  double resid = 0;
  // Blocked so that per-block partial sums add up in a fixed order whatever
  // the thread count.
  const int64_t blk = 4096, nblk = (n + blk - 1) / blk;
  std::vector<double> bres(nblk, 0.0), bdang(nblk, 0.0);
  while (it < max_iter) {
    do_all(pool, 0, nblk, 1, [&](int64_t b, int) {
      double dg = 0;
      for (int64_t u = b * blk; u < std::min(n, (b + 1) * blk); ++u) {
        const uint32_t d = outdeg[u];
        if (d == 0) dg += x[u];
        y[u] = d ? x[u] / d : 0.0;
      }
      bdang[b] = dg;
    });
    double dang = 0;
    for (double d : bdang) dang += d;
    const double base = (1.0 - alpha) / n + alpha * dang / n;
    do_all(pool, 0, nblk, 1, [&](int64_t b, int) {
      double r = 0;
      for (int64_t v = b * blk; v < std::min(n, (b + 1) * blk); ++v) {
        double s = 0;
        for (int32_t e = trow[v]; e < trow[v + 1]; ++e) s += y[tcol[e]];
        xn[v] = base + alpha * s;
        r += std::fabs(xn[v] - x[v]);
      }
      bres[b] = r;
    });
    resid = 0;
    for (double r : bres) resid += r;
    x.swap(xn);
    ++it;
    if (resid < tol) break;
  }
  std::copy(x.begin(), x.end(), x_out);
  if (iters_out) *iters_out = it;
  if (resid_out) *resid_out = resid;
  return 0;
}

int oracle_pagerank_fp64(int64_t n, const int32_t* row, const int32_t* col, double alpha, double tol,
                         uint32_t max_iter, double* x_out, uint32_t* iters_out, double* resid_out, int threads) {
  std::vector<int32_t> trow, tcol;
  transpose(n, row, col, trow, tcol, nullptr);
  std::vector<uint32_t> outdeg(n);
  for (int64_t u = 0; u < n; ++u) outdeg[u] = static_cast<uint32_t>(row[u + 1] - row[u]);
  return oracle_pagerank_fp64_csc(n, outdeg.data(), trow.data(), tcol.data(), alpha, tol, max_iter, x_out,
                                  iters_out, resid_out, threads);
}

// Pull-residual PageRank in FP32, computePRResidual
// (lonestar/analytics/cpu/pagerank/PageRank-pull.cpp:132-190) with
// initNodeDataResidual (:74-88) and computeOutDeg (:90-122), on the original
// CSR (the transpose is built here, in-edges by ascending source).  Both
// sweeps are race free, so with the same per-vertex edge order and FP32
// operation order the result is bit-reproducible.  Returns the iteration
// count.  Pinned by the unnormalised closed form (spsolve of
// x = (1-a) + a*M^T x, to the tolerance the algorithm guarantees).
int oracle_pagerank_residual_f32(int64_t n, const int32_t* row, const int32_t* col, float alpha, float tol,
                                 uint32_t max_iter, float* value_out, uint32_t* iters_out, int threads) {
  Pool pool(threads);
  std::vector<int32_t> trow, tcol;
  transpose(n, row, col, trow, tcol, nullptr);
  std::vector<float> value(n, 0.0f), delta(n, 0.0f), residual(n, 1.0f - alpha);
  std::vector<uint32_t> nout(n);
  for (int64_t u = 0; u < n; ++u) nout[u] = static_cast<uint32_t>(row[u + 1] - row[u]);
  uint32_t it = 0;
  for (;;) {
    std::atomic<uint32_t> accum{0};
    do_all(pool, 0, n, 1024, [&](int64_t s, int) {
      delta[s] = 0;
      if (residual[s] > tol) {
        float old = residual[s];
        residual[s] = 0.0f;
        value[s] += old;
        if (nout[s] > 0) {
          delta[s] = old * alpha / static_cast<float>(nout[s]);
          accum.fetch_add(1, std::memory_order_relaxed);
        }
      }
    });
    do_all(pool, 0, n, 32, [&](int64_t s, int) {
      float sum = 0;
      for (int32_t e = trow[s]; e < trow[s + 1]; ++e) {
        int32_t d = tcol[e];
        if (delta[d] > 0) sum += delta[d];
      }
      if (sum > 0) residual[s] = sum;
    });
    ++it;
    if (it >= max_iter || accum.load() == 0) break;
  }
  std::copy(value.begin(), value.end(), value_out);
  if (iters_out) *iters_out = it;
  return 0;
}

// Jacobi form of computePRTopological (PageRank-pull.cpp:196-251) in FP32:
// value starts 1/n, x' = alpha*sum x[u]/nout[u] + (1-alpha)/n, stop when
// sum|x'-x| <= tol or max_iter.  (The reference updates in place while other
// threads read, so its float results are schedule dependent; this Jacobi
// form is the deterministic member of that family.)
int oracle_pagerank_topo_f32(int64_t n, const int32_t* row, const int32_t* col, float alpha, float tol,
                             uint32_t max_iter, float* value_out, uint32_t* iters_out, int threads) {
  Pool pool(threads);
  std::vector<int32_t> trow, tcol;
  transpose(n, row, col, trow, tcol, nullptr);
  std::vector<float> x(n, 1.0f / static_cast<float>(n)), xn(n);
  std::vector<uint32_t> nout(n);
  for (int64_t u = 0; u < n; ++u) nout[u] = static_cast<uint32_t>(row[u + 1] - row[u]);
  const float base = (1.0f - alpha) / static_cast<float>(n);
  uint32_t it = 0;
  for (;;) {
    double diff = 0;
    std::mutex mu;
    do_all(pool, 0, n, 4096, [&](int64_t v, int) {
      float sum = 0;
      for (int32_t e = trow[v]; e < trow[v + 1]; ++e) {
        int32_t u = tcol[e];
        sum += x[u] / static_cast<float>(nout[u]);
      }
      xn[v] = sum * alpha + base;
    });
    for (int64_t v = 0; v < n; ++v) diff += std::fabs(xn[v] - x[v]);
    x.swap(xn);
    ++it;
    if (diff <= tol || it >= max_iter) break;
  }
  std::copy(x.begin(), x.end(), value_out);
  if (iters_out) *iters_out = it;
  return 0;
}

// Synchronous residual push, PageRank-push.cpp:103-188 (value 0, residual
// 1-alpha, LNode::init :51-56); each round moves residual > tol into value
// and pushes alpha*residual/outdeg to out-neighbours; a neighbour whose
// residual crosses tol joins the next active set.  Serial FP32 (the
// reference's float CAS additions in a fixed order).  Returns rounds.
int oracle_pagerank_push_f32(int64_t n, const int32_t* row, const int32_t* col, float alpha, float tol,
                             uint32_t max_iter, float* value_out, uint32_t* iters_out) {
  std::vector<float> value(n, 0.0f), residual(n, 1.0f - alpha);
  std::vector<uint32_t> active(n);
  for (int64_t v = 0; v < n; ++v) active[v] = static_cast<uint32_t>(v);
  std::vector<char> in_next(n, 0);
  uint32_t it = 0;
  struct Up {
    float delta;
    uint32_t u;
  };
  std::vector<Up> updates;
  for (; !active.empty() && it < max_iter; ++it) {
    updates.clear();
    for (uint32_t s : active) {
      if (residual[s] > tol) {
        float old = residual[s];
        value[s] += old;
        residual[s] = 0.0f;
        int nout = row[s + 1] - row[s];
        float delta = old * alpha / static_cast<float>(nout);
        if (nout > 0) updates.push_back({delta, s});
      }
    }
    std::vector<uint32_t> next;
    for (const Up& up : updates)
      for (int32_t e = row[up.u]; e < row[up.u + 1]; ++e) {
        uint32_t d = static_cast<uint32_t>(col[e]);
        float old = residual[d];
        residual[d] = old + up.delta;
        if (old <= tol && old + up.delta >= tol) next.push_back(d);
      }
    active.swap(next);
  }
  std::copy(value.begin(), value.end(), value_out);
  if (iters_out) *iters_out = it;
  return 0;
}

// Differentiable PageRank, FP64: P[u,v] = w_uv / sum_v' w_uv', dangling rows
// uniform, x = alpha*P^T x + alpha*(D/n) + (1-alpha)*t with D the dangling
// mass; Jacobi to sum|x'-x| < tol or max_iter.  The O(m) part runs in
// parallel as a pull over the in-edges, whose sources are ascending, so every
// x'[v] adds its contributions in the order of the serial push over u: the
// result does not depend on the thread count.  Scalar sums (dangling mass,
// residual) run serially in vertex order.  Pinned by a dense numpy solve on
// small graphs.
int oracle_pagerank_weighted_fp64_mt(int64_t n, const int32_t* row, const int32_t* col, const double* w,
                                     const double* t, double alpha, double tol, uint32_t max_iter, double* x_out,
                                     uint32_t* iters_out, int threads) {
  Pool pool(threads);
  std::vector<double> rs(n, 0.0);
  do_all(pool, 0, n, 1024, [&](int64_t u, int) {
    double s = 0;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) s += w[e];
    rs[u] = s;
  });
  std::vector<int32_t> trow, tcol, perm;
  transpose(n, row, col, trow, tcol, &perm);
  std::vector<double> x(n, 1.0 / n), xn(n), y(n);
  uint32_t it = 0;
  while (it < max_iter) {
    double dang = 0;
    for (int64_t u = 0; u < n; ++u)
      if (row[u + 1] == row[u]) dang += x[u];
    do_all(pool, 0, n, 1024, [&](int64_t v, int) {
      double a = (1.0 - alpha) * t[v] + alpha * dang / n;
      for (int32_t k = trow[v]; k < trow[v + 1]; ++k) {
        const int32_t u = tcol[k];
        a += alpha * x[u] * w[perm[k]] / rs[u];
      }
      xn[v] = a;
    });
    double r = 0;
    for (int64_t v = 0; v < n; ++v) r += std::fabs(xn[v] - x[v]);
    x.swap(xn);
    ++it;
    if (r < tol) break;
  }
  std::copy(x.begin(), x.end(), x_out);
  if (iters_out) *iters_out = it;
  return 0;
}

int oracle_pagerank_weighted_fp64(int64_t n, const int32_t* row, const int32_t* col, const double* w,
                                  const double* t, double alpha, double tol, uint32_t max_iter, double* x_out,
                                  uint32_t* iters_out) {
  return oracle_pagerank_weighted_fp64_mt(n, row, col, w, t, alpha, tol, max_iter, x_out, iters_out, 1);
}

// Adjoint of oracle_pagerank_weighted_fp64 for L = c.x: lambda solves
// lambda = alpha*(P lambda + (1/n) sum(lambda) on dangling rows) + c;
// dL/dt = (1-alpha)*lambda; dL/dw_uv = alpha*x_u*(lambda_v - (P lambda)_u)/s_u
// for a non-dangling row u with weight sum s_u.  Rows are summed in edge
// order in parallel; scalar sums run serially in vertex order, so the result
// does not depend on the thread count.  Pinned by central finite differences
// of oracle_pagerank_weighted_fp64 in tests.
int oracle_pagerank_adjoint_fp64_mt(int64_t n, const int32_t* row, const int32_t* col, const double* w,
                                    const double* x, const double* c, double alpha, double tol, uint32_t max_iter,
                                    double* grad_w, double* grad_t, int threads) {
  Pool pool(threads);
  std::vector<double> rs(n, 0.0);
  do_all(pool, 0, n, 1024, [&](int64_t u, int) {
    double s = 0;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) s += w[e];
    rs[u] = s;
  });
  std::vector<double> lam(c, c + n), ln(n), plam(n);
  for (uint32_t it = 0; it < max_iter; ++it) {
    double tot = 0;
    for (int64_t v = 0; v < n; ++v) tot += lam[v];
    do_all(pool, 0, n, 1024, [&](int64_t u, int) {
      double s;
      if (row[u + 1] == row[u]) {
        s = tot / n;
      } else {
        s = 0;
        for (int32_t e = row[u]; e < row[u + 1]; ++e) s += w[e] * lam[col[e]];
        s /= rs[u];
      }
      ln[u] = c[u] + alpha * s;
    });
    double r = 0;
    for (int64_t u = 0; u < n; ++u) r += std::fabs(ln[u] - lam[u]);
    lam.swap(ln);
    if (r < tol) break;
  }
  do_all(pool, 0, n, 1024, [&](int64_t u, int) {
    double s = 0;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) s += w[e] * lam[col[e]];
    plam[u] = (row[u + 1] > row[u]) ? s / rs[u] : 0.0;
    for (int32_t e = row[u]; e < row[u + 1]; ++e) grad_w[e] = alpha * x[u] * (lam[col[e]] - plam[u]) / rs[u];
  });
  for (int64_t v = 0; v < n; ++v) grad_t[v] = (1.0 - alpha) * lam[v];
  return 0;
}

int oracle_pagerank_adjoint_fp64(int64_t n, const int32_t* row, const int32_t* col, const double* w,
                                 const double* x, const double* c, double alpha, double tol, uint32_t max_iter,
                                 double* grad_w, double* grad_t) {
  return oracle_pagerank_adjoint_fp64_mt(n, row, col, w, x, c, alpha, tol, max_iter, grad_w, grad_t, 1);
}

// ------------------------------------------------------------------- BC
// Brandes betweenness centrality in FP64, BCOuter::doBC
// (lonestar/analytics/cpu/betweennesscentrality/OuterStructs.h:41-102): per
// source a BFS computing sigma and successor lists, then dependencies
// delta[v] = sum_{w in succ(v)} sigma[v]/sigma[w]*(1+delta[w]) in reverse BFS
// order, source excluded; bc[v] += delta[v].  Sources are split statically
// over threads (source i on thread i % T) and per-thread partials are summed
// in thread order, so the result is deterministic for a given thread count.
// Pinned by networkx.betweenness_centrality(normalized=False) on directed
// views of the fixture graphs and brute-force path enumeration.
int oracle_bc(int64_t n, const int32_t* row, const int32_t* col, const uint32_t* sources, uint32_t count,
              double* bc_out, int threads) {
  for (uint32_t i = 0; i < count; ++i)
    if (sources[i] >= n) return 1;
  Pool pool(threads);
  const int T = pool.size();
  std::vector<std::vector<double>> part(T, std::vector<double>(n, 0.0));
  pool.on_each([&](int tid) {
    std::vector<double> sigma(n, 0.0), delta(n, 0.0);
    std::vector<int32_t> d(n, 0);
    std::vector<uint32_t> order;
    order.reserve(n);
    auto& acc = part[tid];
    for (uint32_t i = tid; i < count; i += T) {
      uint32_t s = sources[i];
      order.clear();
      sigma[s] = 1;
      d[s] = 1;
      order.push_back(s);
      for (size_t q = 0; q < order.size(); ++q) {
        uint32_t u = order[q];
        for (int32_t e = row[u]; e < row[u + 1]; ++e) {
          uint32_t v = static_cast<uint32_t>(col[e]);
          if (!d[v]) {
            order.push_back(v);
            d[v] = d[u] + 1;
          }
          if (d[v] == d[u] + 1) sigma[v] += sigma[u];
        }
      }
      for (size_t q = order.size(); q-- > 1;) {
        uint32_t u = order[q];
        double du = 0;
        for (int32_t e = row[u]; e < row[u + 1]; ++e) {
          uint32_t v = static_cast<uint32_t>(col[e]);
          if (d[v] == d[u] + 1) du += (sigma[u] / sigma[v]) * (1.0 + delta[v]);
        }
        delta[u] = du;
      }
      for (uint32_t u : order) {
        acc[u] += delta[u];
        delta[u] = 0;
        sigma[u] = 0;
        d[u] = 0;
      }
    }
  });
  for (int64_t v = 0; v < n; ++v) {
    double s = 0;
    for (int t = 0; t < T; ++t) s += part[t][v];
    bc_out[v] = s;
  }
  return 0;
}

// CSC of a CSR with ascending sources per column (graph-convert transpose),
// exported so tests can compare the library's device transpose.
int oracle_transpose(int64_t n, const int32_t* row, const int32_t* col, int32_t* trow, int32_t* tcol,
                     int32_t* perm) {
  std::vector<int32_t> r, c, p;
  transpose(n, row, col, r, c, &p);
  std::copy(r.begin(), r.end(), trow);
  std::copy(c.begin(), c.end(), tcol);
  std::copy(p.begin(), p.end(), perm);
  return 0;
}

}  // extern "C"
