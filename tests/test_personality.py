"""D-IrGL personality partition (gx_part_*, galois_amd.personality): the
reference's D-IrGL host loops for one host (bfs_push.cpp, sssp_push.cpp,
cc_push.cpp, pagerank_push.cpp, bc_level.cpp, kcore_push.cpp) driven through
the C ABI,
against the oracle and the golden vectors: BFS and SSSP distances bit-exact
(with D-IrGL's infinity), CC labels and k-core membership bit-exact,
PageRank values within the FP32-atomics tolerance, BC within the FP32
dependency tolerance and path counts exact, sanity-check reductions,
split ranges, change bitsets and field accessors."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import csr_arrays

pytestmark = pytest.mark.gpu

ORACLE_INF = 2147483646  # BFS_SSSP.h:31


def _part(row, col, **kw):
    from galois_amd.personality import Partition
    return Partition(np.asarray(row, np.uint32), np.asarray(col, np.uint32), **kw)


def _expected_dist(levels):
    from galois_amd.personality import INFINITY
    lv = np.asarray(levels, np.uint32)
    return np.where(lv == ORACLE_INF, INFINITY, lv).astype(np.uint32)


def test_bfs_push_golden(gx, golden):
    from galois_amd.personality import RANGE_MASTERS, dirgl_bfs, INFINITY
    for g in golden["undirected"]:
        row, col, _ = csr_arrays(g)
        P = _part(row, col)
        for s, exp in g["bfs"].items():
            print(f"dirgl bfs {g['name']} src={s}", flush=True)
            dist, rounds, work = dirgl_bfs(P, int(s))
            np.testing.assert_array_equal(dist, _expected_dist(exp), err_msg=g["name"])
            reached, mx = P.bfs_check(*P.range(RANGE_MASTERS), INFINITY)
            e = np.asarray(exp)
            assert reached == int((e != ORACLE_INF).sum()) and mx == int(e[e != ORACLE_INF].max())


@pytest.mark.parametrize("delta", [0, 2])
def test_bfs_push_rmat(gx, oracle, rmat16, delta):
    """RMAT-16 from the max-degree vertex; delta > 0 defers vertices above the
    round's priority (bfs_push.cpp:219-221) without changing the result."""
    from galois_amd.personality import dirgl_bfs
    G, row, col, _ = rmat16
    s = int(np.argmax(np.diff(row)))
    print(f"dirgl bfs rmat16 delta={delta}", flush=True)
    P = _part(row, col)
    dist, rounds, work = dirgl_bfs(P, s, delta=delta)
    exp = oracle.bfs(row, col, s, threads=8)
    np.testing.assert_array_equal(dist, _expected_dist(exp))
    lv = exp[exp != ORACLE_INF]
    assert rounds >= int(lv.max()) + 1
    # every reached vertex with edges pushes once per lowering; at least its degree
    assert work >= int(np.diff(row)[exp != ORACLE_INF].sum()) - int(np.diff(row)[s])


def test_bfs_push_split_ranges_and_bitset(gx, oracle, rmat16):
    """The operators over three sub-ranges per round give the full-range
    result; FirstItr marks exactly the source's out-neighbours in the
    dist_current bitset, and a range reset clears only that range."""
    from galois_amd.personality import (DIST_CURRENT, DIST_OLD, INFINITY, RANGE_ALL, RANGE_WITH_EDGES)
    G, row, col, _ = rmat16
    s = int(np.argmax(np.diff(row)))
    print("dirgl bfs split ranges", flush=True)
    P = _part(row, col)
    a0, a1 = P.range(RANGE_ALL)
    w0, w1 = P.range(RANGE_WITH_EDGES)
    P.bfs_init(a0, a1, INFINITY, s)
    assert P.get(DIST_CURRENT, s) == 0 and P.get(DIST_OLD, s) == 0
    P.bfs_first(s, s + 1)
    words = P.bitset(DIST_CURRENT)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:P.n].astype(bool)
    nbrs = np.zeros(P.n, bool)
    nbrs[col[row[s]:row[s + 1]]] = True
    np.testing.assert_array_equal(bits, nbrs)
    lo, hi = 1000, 40000
    P.bitset_reset(DIST_CURRENT, lo, hi)
    bits2 = np.unpackbits(P.bitset(DIST_CURRENT).view(np.uint8), bitorder="little")[:P.n].astype(bool)
    keep = nbrs.copy()
    keep[lo:hi] = False
    np.testing.assert_array_equal(bits2, keep)
    cuts = [w0, (w0 + w1) // 3, 2 * (w0 + w1) // 3, w1]
    for _ in range(1000):
        active = sum(P.bfs(cuts[i], cuts[i + 1], 0xFFFFFFFF)[0] for i in range(3))
        if not active:
            break
    assert not active
    np.testing.assert_array_equal(P.read(DIST_CURRENT), _expected_dist(oracle.bfs(row, col, s, threads=8)))


def test_sssp_push(gx, golden, oracle, rmat16):
    """SSSP push distances bit-exact against the golden vectors and the
    oracle (RMAT-16, delta 0 and 5: the priority only defers nodes); the
    sanity check's count, distance sum and maximum."""
    from galois_amd.personality import INFINITY, RANGE_MASTERS, dirgl_sssp
    cases = []
    for g in golden["undirected"]:
        row, col, w = csr_arrays(g)
        for s_, exp in g["sssp"].items():
            cases.append((g["name"], row, col, w, int(s_), np.asarray(exp, np.uint32), 0))
    G, row, col, w = rmat16
    for s_ in (0, int(np.argmax(np.diff(row)))):
        exp = oracle.sssp(row, col, w, s_, 6, threads=8)
        for delta in (0, 5):
            cases.append((f"rmat16 src={s_}", row, col, w, s_, exp, delta))
    for name, row, col, w, src, exp, delta in cases:
        print(f"dirgl sssp {name} src={src} delta={delta}", flush=True)
        P = _part(row, col, weights=w)
        dist, rounds, work = dirgl_sssp(P, src, delta=delta)
        e = _expected_dist(exp)
        np.testing.assert_array_equal(dist, e, err_msg=name)
        reached = e[e != INFINITY].astype(np.int64)
        assert P.sssp_check(*P.range(RANGE_MASTERS), INFINITY) == (len(reached), int(reached.sum()), int(reached.max()))


def test_cc_push(gx, golden, oracle, rmat16):
    """CC push labels (minimum global id of the component) bit-exact against
    the golden vectors and the oracle; the sanity check counts components."""
    from galois_amd.personality import dirgl_cc
    cases = [(g["name"], *csr_arrays(g)[:2], np.asarray(g["cc"], np.uint32)) for g in golden["undirected"]]
    G, row, col, _ = rmat16
    cases.append(("rmat16", row, col, oracle.cc(row, col, threads=8)))
    for name, row, col, exp in cases:
        print(f"dirgl cc {name}", flush=True)
        P = _part(row, col)
        label, comps, rounds = dirgl_cc(P)
        np.testing.assert_array_equal(label, exp, err_msg=name)
        assert comps == int((exp == np.arange(len(exp))).sum())


def test_pagerank_push(gx, golden, oracle):
    """PageRank push values (unnormalised, reset probability 0.15) against the
    golden closed form and the oracle's FP32 residual pull (the same
    unnormalised fixed point), within the tolerance the FP32 atomics allow; PageRankSanity's reductions against the
    fields read back.  InitializeGraph and PageRank_delta run over the nodes
    with edges only (pagerank_push.cpp:127-128, 189), which D-IrGL partitions
    number first: nodes past the last node with out-edges never receive a
    residual and keep value 0."""
    from galois_amd.personality import PR_RESIDUAL, RANGE_MASTERS, RANGE_WITH_EDGES, dirgl_pagerank
    tol = 1e-7
    for g in golden["undirected"] + golden["directed"]:
        print(f"dirgl pagerank {g['name']}", flush=True)
        row, col, _ = csr_arrays(g)
        P = _part(row, col)
        value, rounds = dirgl_pagerank(P, tol)
        we = P.range(RANGE_WITH_EDGES)[1]
        closed = np.asarray(g["pagerank_unnormalised"]).copy()
        closed[we:] = 0.0
        np.testing.assert_allclose(value, closed, atol=2e-5, err_msg=g["name"])
        ref = oracle.pagerank_residual(row, col, tol=tol)[0].astype(np.float64)
        ref[we:] = 0.0
        np.testing.assert_allclose(value, ref, atol=2e-5, err_msg=g["name"])
        san = P.pr_check(*P.range(RANGE_MASTERS), tol)
        res = P.read_float(PR_RESIDUAL)
        assert san.max_value == value.max() and san.min_value == value.min()
        assert san.max_residual == res.max() and san.min_residual == res.min()
        assert abs(san.sum_value - float(value.astype(np.float64).sum())) < 1e-4
        assert san.residual_over_tolerance == int((res > np.float32(tol)).sum())


def test_bc_level(gx, golden, oracle, rmat16):
    """BC level against the oracle's FP64 Brandes BC (OuterStructs.h:41-102)
    on the golden graphs (all sources) and RMAT-16 (16 seeded sources): FP32
    dependencies, 1e-4 relative; path counts of
    the last source against a BFS path count; Sanity's reductions."""
    from galois_amd.personality import BC_LENGTH, RANGE_MASTERS, dirgl_bc
    cases = [(g["name"], *csr_arrays(g)[:2], list(range(g["n"]))) for g in golden["undirected"]]
    G, row, col, _ = rmat16
    cases.append(("rmat16", row, col, [int(x) for x in gx.pick_sources(row, 16)]))
    for name, row, col, srcs in cases:
        print(f"dirgl bc {name} sources={len(srcs)}", flush=True)
        P = _part(row, col)
        bc = dirgl_bc(P, srcs)
        exp = oracle.bc(row, col, np.asarray(srcs, np.uint32), threads=8)
        scale = np.maximum(np.abs(exp), 1.0)
        assert np.all(np.abs(bc - exp) <= 1e-4 * scale), (name, float(np.max(np.abs(bc - exp) / scale)))
        # shortest-path counts of the last source: sigma[v] = sum over BFS parents
        lv = oracle.bfs(row, col, srcs[-1], threads=8).astype(np.int64)
        sigma = np.zeros(len(lv))
        sigma[srcs[-1]] = 1.0
        order = np.argsort(lv, kind="stable")
        rr = np.asarray(row, np.int64)
        for u in order:
            if lv[u] == ORACLE_INF:
                break
            for v in col[rr[u]:rr[u + 1]]:
                if lv[v] == lv[u] + 1:
                    sigma[v] += sigma[u]
        np.testing.assert_array_equal(P.read_paths(), sigma, err_msg=name)
        s, mx, mn = P.bc_check(*P.range(RANGE_MASTERS))
        assert mx == bc.max() and mn == bc.min() and abs(s - float(bc.astype(np.float64).sum())) <= 1e-5 * max(1.0, s)


def test_kcore_push(gx, golden, oracle, rmat16):
    """k-core membership (flag) against the oracle's peeling (alive iff final
    degree >= k) and the golden vectors; nodes past the last node with
    edges are never visited by KCoreStep1 and keep their flag, as in the
    reference; KCoreSanityCheck counts the masters with the flag set."""
    from galois_amd.personality import CURRENT_DEGREE, dirgl_kcore
    cases = [(g["name"], *csr_arrays(g)[:2], int(k), np.asarray(e["alive"], bool))
             for g in golden["undirected"] for k, e in g["kcore"].items()]
    G, row, col, _ = rmat16
    for k in (2, 10, 40):
        cases.append((f"rmat16 k={k}", row, col, k, oracle.kcore(row, col, k, threads=8) >= k))
    for name, row, col, k, alive in cases:
        print(f"dirgl kcore {name} k={k}", flush=True)
        P = _part(row, col)
        flag, count, rounds = dirgl_kcore(P, k)
        deg = np.diff(np.asarray(row, np.int64))
        nz = np.nonzero(deg)[0]
        with_edges = int(nz[-1]) + 1 if len(nz) else 0
        exp = alive.copy()
        exp[with_edges:] = True
        np.testing.assert_array_equal(flag.astype(bool), exp, err_msg=name)
        assert count == int(exp.sum())
        # survivors keep a degree of at least k counted over live neighbours
        cur = P.read(CURRENT_DEGREE)
        assert np.all(cur[:with_edges][exp[:with_edges]] >= k), name


def test_field_accessors_and_validation(gx):
    from galois_amd._lib import GxError
    from galois_amd.personality import DIST_CURRENT, RANGE_MASTERS, TRIM
    row = np.array([0, 2, 3, 3, 4], np.uint32)
    col = np.array([1, 2, 3, 0], np.uint32)
    P = _part(row, col, num_owned=3, begin_master=1)
    assert P.range(RANGE_MASTERS) == (1, 4)
    P.fill(TRIM, 0, 4, 7)
    P.add(TRIM, 2, 5)
    assert list(P.read(TRIM)) == [7, 7, 12, 7]
    P.set(DIST_CURRENT, 3, 9)
    assert P.min(DIST_CURRENT, 3, 4) and not P.min(DIST_CURRENT, 3, 6)
    assert P.get(DIST_CURRENT, 3) == 4
    with pytest.raises(GxError):
        P.get(DIST_CURRENT, 4)  # outside the partition
    with pytest.raises(GxError):
        P.bfs(0, 5, 1)
    with pytest.raises(GxError):
        _part(row, np.array([1, 2, 9, 0], np.uint32))  # destination out of range
    with pytest.raises(GxError):
        _part(np.array([0, 2, 1, 3, 4], np.uint32), col)  # offsets not monotone
