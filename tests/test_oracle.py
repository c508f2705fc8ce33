"""Pins the oracle (CPU restatement of the reference) against golden vectors,
closed forms and invariants.  CPU only.

Each oracle function is checked by something that would fail on a plausible
mistake (wrong level/offset, dropped dangling term, wrong tie-break of the
union-find, transposed operand, missing 1+delta term ...):
  BFS      networkx levels + BFS_SSSP::verify invariants + tightness
  SSSP     scipy dijkstra + the serial Dijkstra restatement
  CC       scipy components relabelled to component minima
  k-core   networkx.k_core + alive-neighbour degree identity
  PR fp64  dense direct solve of the normalised system
  PR ref   dense solve of the unnormalised system (to its tolerance)
  BC       networkx (all sources) and brute-force path enumeration (subset)
  adjoint  finite differences of a dense solve
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import csr_arrays

INF = 2147483646


def _cases(golden):
    return golden["undirected"]


def test_bfs_matches_networkx(oracle, golden):
    for g in _cases(golden):
        row, col, _ = csr_arrays(g)
        for s, exp in g["bfs"].items():
            got = oracle.bfs(row, col, int(s))
            np.testing.assert_array_equal(got, np.asarray(exp, np.uint32), err_msg=g["name"])
    for g in golden["directed"]:
        row, col, _ = csr_arrays(g)
        np.testing.assert_array_equal(oracle.bfs(row, col, 0), np.asarray(g["bfs0"], np.uint32))


def test_bfs_invariants_random(oracle):
    rng = np.random.default_rng(1)
    n = 2000
    u = rng.integers(0, n, 12000)
    v = rng.integers(0, n, 12000)
    keys = np.unique(np.concatenate([u * n + v, v * n + u]))
    keys = keys[(keys // n) != (keys % n)]
    src, dst = keys // n, keys % n
    row = np.zeros(n + 1, np.int32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    col = dst.astype(np.int32)
    lv = oracle.bfs(row, col, 0, threads=8).astype(np.int64)
    assert lv[0] == 0
    reached = lv != INF
    # every edge: level[dst] <= level[src] + 1 (BFS_SSSP.h:231-263)
    ok = ~reached[src] | (lv[dst] <= lv[src] + 1)
    assert ok.all()
    # tightness: every reached v != src has a neighbour one level closer
    for x in np.flatnonzero(reached)[1:200]:
        nb = col[row[x]:row[x + 1]]
        assert (lv[nb] == lv[x] - 1).any()


def test_sssp_matches_scipy_and_dijkstra(oracle, golden):
    for g in _cases(golden):
        row, col, w = csr_arrays(g)
        for s, exp in g["sssp"].items():
            exp = np.asarray(exp, np.uint32)
            for shift in (0, 2, 6, 13):
                np.testing.assert_array_equal(oracle.sssp(row, col, w, int(s), shift), exp, err_msg=g["name"])
            np.testing.assert_array_equal(oracle.dijkstra(row, col, w, int(s)), exp)


def test_cc_min_labels(oracle, golden):
    for g in _cases(golden):
        row, col, _ = csr_arrays(g)
        np.testing.assert_array_equal(oracle.cc(row, col), np.asarray(g["cc"], np.uint32), err_msg=g["name"])


def test_kcore_matches_networkx(oracle, golden):
    for g in _cases(golden):
        row, col, _ = csr_arrays(g)
        for k, exp in g["kcore"].items():
            deg = oracle.kcore(row, col, int(k))
            alive = (deg >= int(k)).astype(np.int64)
            np.testing.assert_array_equal(alive, exp["alive"], err_msg=f"{g['name']} k={k}")
            np.testing.assert_array_equal(deg, np.asarray(exp["degree"], np.uint32))


def test_pagerank_fp64_closed_form(oracle, golden):
    for g in _cases(golden) + golden["directed"]:
        row, col, _ = csr_arrays(g)
        x, it, res = oracle.pagerank(row, col, tol=1e-14, max_iter=2000)
        exp = np.asarray(g["pagerank"])
        assert abs(x.sum() - 1.0) < 1e-12
        np.testing.assert_allclose(x, exp, rtol=1e-10, atol=1e-13, err_msg=g["name"])
        assert res < 1e-14


def test_pagerank_reference_variants_closed_form(oracle, golden):
    # Residual/topological pull and sync push converge to the unnormalised
    # fixed point x = (1-a) + a*M^T x (dangling mass dropped) within a bound
    # proportional to the tolerance.
    for g in _cases(golden) + golden["directed"]:
        row, col, _ = csr_arrays(g)
        exp = np.asarray(g["pagerank_unnormalised"])
        xr, _ = oracle.pagerank_residual(row, col, tol=1e-7)
        np.testing.assert_allclose(xr, exp, atol=2e-5, err_msg=g["name"])
        xp, _ = oracle.pagerank_push(row, col, tol=1e-7)
        np.testing.assert_allclose(xp, exp, atol=2e-5, err_msg=g["name"])
        n = len(row) - 1
        xt, _ = oracle.pagerank_topo(row, col, tol=1e-7)
        # topological starts from 1/n with base (1-a)/n: its fixed point is exp/n
        np.testing.assert_allclose(xt, exp / n, atol=1e-6, err_msg=g["name"])


def test_bc_networkx_and_bruteforce(oracle, golden):
    for g in _cases(golden):
        row, col, _ = csr_arrays(g)
        n = len(row) - 1
        allbc = oracle.bc(row, col, np.arange(n))
        np.testing.assert_allclose(allbc, g["bc_all"], rtol=1e-12, atol=1e-12, err_msg=g["name"])
        sub = oracle.bc(row, col, g["bc_sources"])
        np.testing.assert_allclose(sub, g["bc_subset"], rtol=1e-12, atol=1e-12, err_msg=g["name"])


def test_bc_dependency_sum_invariant(oracle, golden):
    # sum_v delta_s(v) = sum_{t reachable, t != s} (d(s,t) - 1) for one source
    for g in _cases(golden):
        row, col, _ = csr_arrays(g)
        s = g["bc_sources"][0]
        bc = oracle.bc(row, col, [s])
        lv = oracle.bfs(row, col, s).astype(np.int64)
        reach = (lv != INF) & (lv > 0)
        assert abs(bc.sum() - (lv[reach] - 1).sum()) < 1e-9


def test_weighted_pagerank_and_adjoint(oracle, golden):
    for g in golden["weighted"]:
        row = np.asarray(g["row"], np.int32)
        col = np.asarray(g["col"], np.int32)
        x = oracle.pagerank_weighted(row, col, g["w"], g["t"], alpha=g["alpha"])
        np.testing.assert_allclose(x, g["x"], rtol=1e-10)
        gw, gt = oracle.pagerank_adjoint(row, col, g["w"], x, g["c"], alpha=g["alpha"])
        np.testing.assert_allclose(gw, g["grad_w"], rtol=1e-5, atol=1e-8)
        np.testing.assert_allclose(gt, g["grad_t"], rtol=1e-5, atol=1e-8)


def test_weighted_pagerank_threads_bit_identical(oracle):
    """The threaded forward and adjoint (pull over ascending sources, scalar
    sums serial) equal the serial restatement bit for bit, so the fixtures that
    pin the serial one pin them too."""
    rng = np.random.default_rng(7)
    n = 3000
    deg = rng.integers(0, 12, n)
    deg[rng.integers(0, n, 40)] = 0  # dangling rows
    row = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    col = rng.integers(0, n, row[-1]).astype(np.int32)
    col[::97] = col[::97] // 2  # some duplicate targets inside rows
    w = rng.lognormal(0, 1, row[-1])
    t = rng.random(n)
    t /= t.sum()
    c = rng.standard_normal(n)
    x1 = oracle.pagerank_weighted(row, col, w, t, tol=1e-13, max_iter=300)
    x4 = oracle.pagerank_weighted(row, col, w, t, tol=1e-13, max_iter=300, threads=4)
    np.testing.assert_array_equal(x1, x4)
    g1 = oracle.pagerank_adjoint(row, col, w, x1, c, tol=1e-13, max_iter=300)
    g4 = oracle.pagerank_adjoint(row, col, w, x1, c, tol=1e-13, max_iter=300, threads=4)
    np.testing.assert_array_equal(g1[0], g4[0])
    np.testing.assert_array_equal(g1[1], g4[1])


def test_transpose_order(oracle, golden):
    for g in _cases(golden) + golden["directed"]:
        row, col, _ = csr_arrays(g)
        trow, tcol, perm = oracle.transpose(row, col)
        n = len(row) - 1
        src = np.repeat(np.arange(n), np.diff(row))
        # CSC slot j holds edge perm[j]: same endpoints, sources ascending per column
        np.testing.assert_array_equal(tcol, src[perm])
        np.testing.assert_array_equal(col[perm], np.repeat(np.arange(n), np.diff(trow)))
        for v in range(n):
            seg = tcol[trow[v]:trow[v + 1]]
            assert np.all(np.diff(seg) > 0)


def test_checker_detects_planted_mistakes(oracle, golden):
    """A wrong oracle must fail the golden checks above (sensitivity)."""
    g = _cases(golden)[4]  # er40
    row, col, w = csr_arrays(g)
    lv = oracle.bfs(row, col, 0)
    bad = lv.copy()
    bad[np.flatnonzero(bad != INF)[-1]] += 1  # off-by-one level
    assert not np.array_equal(bad, np.asarray(g["bfs"]["0"], np.uint32))
    x, _, _ = oracle.pagerank(row, col, tol=1e-14, max_iter=2000)
    # dropping the dangling term / teleport scaling changes the vector
    assert not np.allclose(x * 1.001, g["pagerank"], rtol=1e-10)
    lab = oracle.cc(row, col)
    # union-find that links the lower root under the higher would give max labels
    comp_max = {}
    for v, l in enumerate(lab):
        comp_max[l] = max(comp_max.get(l, v), v)
    wrong = np.asarray([comp_max[l] for l in lab], np.uint32)
    if not np.array_equal(wrong, lab):
        assert not np.array_equal(wrong, np.asarray(g["cc"], np.uint32))
