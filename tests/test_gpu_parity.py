"""GPU parity: the HIP library through the C ABI against golden vectors and the
oracle on the same inputs.  Integer outputs are bit-exact; PageRank is within
the tolerances stated per test."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import csr_arrays

pytestmark = pytest.mark.gpu

INF = 2147483646


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def dev(gx, g, symmetric=True, weights=True):
    row, col, w = csr_arrays(g)
    return gx.Graph.from_csr(row, col, w if weights else None, symmetric=symmetric)


# ------------------------------------------------------------- golden graphs
def test_bfs_golden(gx, golden):
    for g in golden["undirected"]:
        G = dev(gx, g)
        for s, exp in g["bfs"].items():
            say(f"bfs {g['name']} src={s}")
            lv, st = gx.bfs(G, int(s))
            np.testing.assert_array_equal(u32(lv), np.asarray(exp, np.uint32), err_msg=g["name"])
            lv2, par, _ = gx.bfs_do(G, int(s), parents=True)
            np.testing.assert_array_equal(u32(lv2), np.asarray(exp, np.uint32), err_msg=g["name"])
            _check_parents(g, int(s), u32(lv2), u32(par))
    for g in golden["directed"]:
        G = dev(gx, g, symmetric=False, weights=False)
        lv, _ = gx.bfs(G, 0)
        np.testing.assert_array_equal(u32(lv), np.asarray(g["bfs0"], np.uint32))
        lv2, _, _ = gx.bfs_do(G, 0, alpha=1, beta=1)  # forces pull levels on the CSC
        np.testing.assert_array_equal(u32(lv2), np.asarray(g["bfs0"], np.uint32))


def _check_parents(g, s, lv, par):
    row, col, _ = csr_arrays(g)
    for v in range(len(row) - 1):
        if lv[v] == INF:
            assert par[v] == 0xFFFFFFFF
        elif v == s:
            assert par[v] == s
        else:
            p = par[v]
            assert lv[p] == lv[v] - 1
            assert v in col[row[p]:row[p + 1]]


def test_sssp_golden(gx, golden):
    for g in golden["undirected"]:
        G = dev(gx, g)
        for s, exp in g["sssp"].items():
            for shift in (0, 2, 6, 13):
                say(f"sssp {g['name']} src={s} shift={shift}")
                d, _ = gx.sssp(G, int(s), shift)
                np.testing.assert_array_equal(u32(d), np.asarray(exp, np.uint32), err_msg=g["name"])


def test_cc_golden(gx, golden):
    for g in golden["undirected"]:
        say(f"cc {g['name']}")
        lab, _ = gx.connected_components(dev(gx, g))
        np.testing.assert_array_equal(u32(lab), np.asarray(g["cc"], np.uint32), err_msg=g["name"])


def test_kcore_golden(gx, golden):
    for g in golden["undirected"]:
        G = dev(gx, g)
        for k, exp in g["kcore"].items():
            say(f"kcore {g['name']} k={k}")
            deg, alive, cnt, _ = gx.kcore(G, int(k))
            np.testing.assert_array_equal(u32(deg), np.asarray(exp["degree"], np.uint32))
            np.testing.assert_array_equal(alive.cpu().numpy(), np.asarray(exp["alive"], np.uint8))
            assert cnt == sum(exp["alive"])


def test_pagerank_golden(gx, golden):
    for g in golden["undirected"] + golden["directed"]:
        say(f"pagerank {g['name']}")
        G = dev(gx, g, symmetric=False, weights=False)
        exp = np.asarray(g["pagerank"])
        x64, st = gx.pagerank(G, tol=1e-14, max_iter=3000, precision="fp64")
        np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-10, atol=1e-13, err_msg=g["name"])
        assert st.residual < 1e-14
        x32, _ = gx.pagerank(G, tol=1e-12, max_iter=300, precision="fp32")
        # FP32 (compensated sums) vs FP64 closed form: L1 error bound 1e-6
        assert np.abs(x32.cpu().numpy().astype(np.float64) - exp).sum() < 1e-6, g["name"]


def test_pagerank_reference_semantics(gx, golden, oracle):
    for g in golden["undirected"] + golden["directed"]:
        say(f"pagerank_ref {g['name']}")
        row, col, _ = csr_arrays(g)
        G = dev(gx, g, symmetric=False, weights=False)
        # residual pull is race free: bit-identical to the CPU restatement
        for tol in (1e-3, 1e-6):
            got, st = gx.pagerank_ref(G, "residual", tol=tol)
            exp, it = oracle.pagerank_residual(row, col, tol=tol)
            np.testing.assert_array_equal(got.cpu().numpy(), exp, err_msg=g["name"])
            assert st.iterations == it
        got, _ = gx.pagerank_ref(G, "topo", tol=1e-7)
        exp, _ = oracle.pagerank_topo(row, col, tol=1e-7)
        np.testing.assert_allclose(got.cpu().numpy(), exp, rtol=1e-5, atol=1e-9)
        # push: FP32 atomics (order dependent): within the tolerance bound
        got, _ = gx.pagerank_push(G, tol=1e-7)
        closed = np.asarray(g["pagerank_unnormalised"])
        np.testing.assert_allclose(got.cpu().numpy(), closed, atol=2e-5, err_msg=g["name"])


# ------------------------------------------------------------- generators
def _mix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15))
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _hash2(seed, a, b):
    return _mix64(_mix64(np.uint64(seed) ^ _mix64(a)) ^ (b * np.uint64(0xD1B54A32D192ED03)))


def numpy_rmat(scale, ef, a, b, c, seed, symmetric):
    """Independent numpy statement of the counter-based RMAT generator."""
    with np.errstate(over="ignore"):
        m = (1 << scale) * ef
        i = np.arange(m, dtype=np.uint64)
        u = np.zeros(m, np.uint64)
        v = np.zeros(m, np.uint64)
        ta, tab, tabc = (np.uint64(int(np.floor(p * 2 ** 32))) for p in (a, a + b, a + b + c))
        for lvl in range(0, scale, 2):
            h = _hash2(seed, i, np.uint64(lvl))
            for k in range(2):
                if lvl + k >= scale:
                    break
                r = (h & np.uint64(0xFFFFFFFF)) if k == 0 else (h >> np.uint64(32))
                bit = np.uint64(1 << (scale - 1 - (lvl + k)))
                colb = (r >= ta) & ((r < tab) | (r >= tabc))
                rowb = r >= tab
                v |= np.where(colb, bit, np.uint64(0))
                u |= np.where(rowb, bit, np.uint64(0))
    if symmetric:
        u, v = np.concatenate([u, v]), np.concatenate([v, u])
    keep = u != v
    keys = np.unique(u[keep] * np.uint64(1 << scale) + v[keep])
    src = (keys >> np.uint64(scale)).astype(np.int64)
    dst = (keys & np.uint64((1 << scale) - 1)).astype(np.int32)
    row = np.zeros((1 << scale) + 1, np.int64)
    np.add.at(row, src + 1, 1)
    return np.cumsum(row).astype(np.int32), dst


@pytest.mark.parametrize("symmetric", [True, False])
def test_rmat_generator_matches_numpy_statement(gx, symmetric):
    say(f"rmat generator symmetric={symmetric}")
    G = gx.Graph.rmat(10, 8, seed=12345, symmetric=symmetric)
    row, col, _ = G.to_host()
    erow, ecol = numpy_rmat(10, 8, 0.57, 0.19, 0.19, 12345, symmetric)
    np.testing.assert_array_equal(row, erow)
    np.testing.assert_array_equal(col, ecol)
    # deterministic: same seed, same graph
    row2, col2, _ = gx.Graph.rmat(10, 8, seed=12345, symmetric=symmetric).to_host()
    np.testing.assert_array_equal(col, col2)


def test_transpose_matches_oracle(gx, oracle):
    say("transpose rmat12")
    G = gx.Graph.rmat(12, 16, symmetric=False)
    row, col, _ = G.to_host()
    trow, tcol, perm = G.transpose_host()
    erow, ecol, eperm = oracle.transpose(row, col)
    np.testing.assert_array_equal(trow, erow)
    np.testing.assert_array_equal(tcol, ecol)
    np.testing.assert_array_equal(perm, eperm)


def test_grid_generator_properties(gx):
    say("grid 64")
    G = gx.Graph.grid(64)
    row, col, w = G.to_host()
    n = 64 * 64
    assert len(row) == n + 1
    src = np.repeat(np.arange(n), np.diff(row))
    # symmetric with symmetric weights, no self loops, sorted unique rows
    key = src.astype(np.int64) * n + col
    rkey = col.astype(np.int64) * n + src
    order = np.argsort(rkey)
    np.testing.assert_array_equal(np.sort(rkey), key)
    np.testing.assert_array_equal(w[order], w)
    assert (src != col).all()
    assert (w >= 1).all() and (w <= 1000).all()
    lattice = 2 * 64 * 63
    assert 0.9 * 2 * lattice < len(col) < 1.05 * 2 * lattice


# ------------------------------------------------------------- RMAT parity
def test_rmat16_bfs_sssp(gx, oracle, rmat16):
    G, row, col, w = rmat16
    deg = np.diff(row)
    for s in (int(np.argmax(deg)), 0, 12345):
        say(f"rmat16 bfs/sssp src={s}")
        exp = oracle.bfs(row, col, s, threads=8)
        lv, _ = gx.bfs(G, s)
        np.testing.assert_array_equal(u32(lv), exp)
        lv2, _, _ = gx.bfs_do(G, s)
        np.testing.assert_array_equal(u32(lv2), exp)
        for shift in (4, 6):
            d, _ = gx.sssp(G, s, shift)
            np.testing.assert_array_equal(u32(d), oracle.sssp(row, col, w, s, shift, threads=8))


def test_rmat16_cc_kcore(gx, oracle, rmat16):
    G, row, col, w = rmat16
    say("rmat16 cc")
    lab, _ = gx.connected_components(G)
    np.testing.assert_array_equal(u32(lab), oracle.cc(row, col, threads=8))
    for k in (2, 10, 40):
        say(f"rmat16 kcore k={k}")
        deg, alive, cnt, _ = gx.kcore(G, k)
        exp = oracle.kcore(row, col, k, threads=8)
        np.testing.assert_array_equal(u32(deg), exp)
        assert cnt == int((exp >= k).sum())


def test_rmat16_pagerank(gx, oracle, rmat16):
    G, row, col, w = rmat16
    say("rmat16 pagerank")
    exp, it, res = oracle.pagerank(row, col, tol=1e-9, max_iter=100, threads=8)
    x64, st = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp64")
    assert st.iterations == it
    np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-11, atol=1e-17)
    x32, st32 = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    err = np.abs(x32.cpu().numpy().astype(np.float64) - exp)
    # FP32 stops when its own residual (rounding-order dependent) drops below
    # tol: within a couple of iterations of the FP64 count.
    assert err.sum() < 1e-6 and abs(st32.iterations - it) <= 2


@pytest.mark.parametrize("hot", [0, 700, 20000])
def test_pagerank_hot_cold_split(gx, oracle, rmat16, hot):
    """The two-pass pull (sources below `hot` in degree order summed first)
    gives the single-pass result: FP64 to 1e-11 relative, FP32 to 1e-6 L1."""
    G, row, col, w = rmat16
    say(f"rmat16 pagerank hot={hot}")
    exp, it, _ = oracle.pagerank(row, col, tol=1e-9, max_iter=100, threads=8)
    gx.tune_set("pr.hot_sources", hot)
    gx.tune_set("pr.blocking", 0)  # FP32 through the merge-path pull as well
    try:
        x64, st = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp64")
        x32, st32 = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    finally:
        gx.tune_set("pr.hot_sources", -1)
        gx.tune_set("pr.blocking", -1)
    assert st.iterations == it
    np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-11, atol=1e-17)
    assert np.abs(x32.cpu().numpy().astype(np.float64) - exp).sum() < 1e-6


@pytest.mark.parametrize("scale", [16, 20])
def test_pagerank_blocked_vs_pull(gx, oracle, scale):
    """FP32 propagation-blocked sweep (exact fixed-point row sums) and the
    FP32 merge-path pull (Kahan row sums) against the FP64 oracle: L1 < 1e-6;
    FP32 stops on its own residual, which reaches 1e-9 no earlier than FP64's
    (minus rounding luck) and often later; the blocked result is
    bit-identical across runs.  At scale 20 the heaviest destination bins
    exceed one task and are summed in slices (one extra launch per
    iteration)."""
    say(f"rmat{scale} pagerank blocked vs pull")
    G = gx.Graph.rmat(scale, 16, symmetric=False)
    row, col, _ = G.to_host()
    exp, it, _ = oracle.pagerank(row, col, tol=1e-9, max_iter=100, threads=8)
    xb, stb = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    xb2, _ = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    gx.tune_set("pr.blocking", 0)
    try:
        xp, stp = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    finally:
        gx.tune_set("pr.blocking", -1)
    for x, st in ((xb, stb), (xp, stp)):
        err = np.abs(x.cpu().numpy().astype(np.float64) - exp)
        assert err.sum() < 1e-6 and it - 2 <= st.iterations < 100, (err.sum(), st.iterations, it)
    np.testing.assert_array_equal(xb.cpu().numpy(), xb2.cpu().numpy())
    # stated bytes per sweep: blocked 2m + 10g + 20n + 12/job with g groups
    # (0 < g <= m); pull 4m + 4(n+1) + 20n, more with the hot/cold split
    m, n = int(row[-1]), len(row) - 1
    assert 2 * m + 20 * n < stb.main_kernel_bytes, stb.main_kernel_bytes
    assert stp.main_kernel_bytes >= 4 * m + 24 * n + 4, stp.main_kernel_bytes
    # FP64 through the blocked sweep (knob 2; FP64 atomic row sums): the
    # oracle's iteration count and 1e-11 relative, as the FP64 pull.
    gx.tune_set("pr.blocking", 2)
    try:
        x64, st64 = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp64")
    finally:
        gx.tune_set("pr.blocking", -1)
    assert st64.iterations == it
    np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-11, atol=1e-17)
    if scale == 20:
        # init (3) + unpermute (1) + per enqueued iteration (batches of 8):
        # gather, accumulate, sliced bins, reduce
        enq = min(100, -(-stb.iterations // 8) * 8)
        assert stb.launches == 4 + 4 * enq, (stb.launches, stb.iterations)


@pytest.mark.parametrize("knobs", [{"pr.task_edges": 1 << 12}, {"pr.task_edges": 1 << 30}])
def test_pagerank_task_knobs(gx, knobs):
    """The task size of the blocked sweep only decides how a bin's in-edges
    are split over workgroups (2^12: every bin of RMAT-18 in dozens of slices
    meeting in the global accumulator; 2^30: no bin sliced).  The fixed-point
    row sums are exact, so the FP32 ranks and the iteration count are
    bit-identical to the default's (2^19)."""
    say(f"rmat18 pagerank {knobs}")
    G = gx.Graph.rmat(18, 16, symmetric=False)
    assert gx.pagerank_layout_info(G) == {}
    ref, st_ref = gx.pagerank(G, tol=1e-9, max_iter=100, precision="fp32")
    info_ref = gx.pagerank_layout_info(G)
    assert info_ref["bins"] == 16 and info_ref["task_edges"] == 1 << 19
    try:
        for k, v in knobs.items():
            gx.tune_set(k, v)
        G2 = gx.Graph.rmat(18, 16, symmetric=False)  # its layout is built under the knobs
        x, st = gx.pagerank(G2, tol=1e-9, max_iter=100, precision="fp32")
    finally:
        for k in knobs:
            gx.tune_set(k, -1)
    info = gx.pagerank_layout_info(G2)
    assert info["task_edges"] == knobs.get("pr.task_edges", 1 << 19)
    assert info["groups"] == info_ref["groups"] and info["bins"] == info_ref["bins"]
    if info["task_edges"] < 1 << 13:
        assert info["sliced_bins"] > 0 and info["tasks"] > info_ref["tasks"]
    else:
        assert info["sliced_bins"] == 0 and info["tasks"] == 16
    assert st.iterations == st_ref.iterations
    np.testing.assert_array_equal(x.cpu().numpy(), ref.cpu().numpy())


def test_pagerank_blocked_shapes(gx, oracle):
    """Propagation-blocked sweep on shapes its tables must split: a source
    with 600K out-edges (one chunk in several phase-1 pieces), a destination
    with 600K in-edges (its bin sliced over several tasks plus the global
    accumulator), n one past a 16K chunk boundary (a one-vertex chunk and
    bin), a band of vertices with neither in- nor out-edges (after the degree
    order, empty chunks at the end), dangling vertices.  FP32 vs the FP64 oracle to 1e-6 in L1, and
    deterministic."""
    say("pagerank blocked shapes")
    rng = np.random.default_rng(11)
    n = 5 * 16384 + 16385  # 98,305 vertices: last chunk holds one vertex
    k = 600000
    hub_out, hub_in = 70000, 3
    srcs = [np.full(k, hub_out), rng.integers(0, n, k), rng.integers(0, 16384, 200000)]
    dsts = [rng.integers(0, n, k), np.full(k, hub_in), rng.integers(0, n, 200000)]
    src = np.concatenate(srcs).astype(np.int64)
    dst = np.concatenate(dsts).astype(np.int64)
    quiet = (src >= 32768) & (src < 49152) | (dst >= 32768) & (dst < 49152)  # chunk and bin 2 empty
    src, dst = src[~quiet], dst[~quiet]
    src = np.append(src, n - 1)
    dst = np.append(dst, 0)  # the one-vertex chunk has an out-edge
    keys = np.unique(src * n + dst)
    s, d = keys // n, keys % n
    keep = s != d
    s, d = s[keep], d[keep]
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, s + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    col = d.astype(np.int32)
    G = gx.Graph.from_csr(row, col, None, symmetric=False)
    exp, it, _ = oracle.pagerank(row, col, tol=1e-10, max_iter=200, threads=8)
    x1, st1 = gx.pagerank(G, tol=1e-10, max_iter=200, precision="fp32")
    x2, _ = gx.pagerank(G, tol=1e-10, max_iter=200, precision="fp32")
    a = x1.cpu().numpy()
    assert np.abs(a.astype(np.float64) - exp).sum() < 1e-6
    np.testing.assert_array_equal(a, x2.cpu().numpy())
    gx.tune_set("pr.blocking", 2)
    try:
        x64, st64 = gx.pagerank(G, tol=1e-10, max_iter=200, precision="fp64")
    finally:
        gx.tune_set("pr.blocking", -1)
    assert st64.iterations == it
    np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-11, atol=1e-18)
    # tol 1e-10 is below FP32 resolution here: most x values stop changing and
    # FP32's own residual can reach tol before FP64's does; accuracy is the
    # check above.
    assert st1.iterations <= 200 and (st1.residual < 1e-10 or st1.iterations == 200)


@pytest.mark.parametrize("vis", [0, 1])
def test_bfs_visited_bitmap_knob(gx, oracle, rmat16, vis):
    """Top-down BFS with and without the visited bitmap (used by default from
    2^25 vertices on): levels bit-exact vs the oracle either way."""
    G, row, col, _ = rmat16
    say(f"rmat16 bfs visited_bitmap={vis}")
    src = int(np.argmax(np.diff(row)))
    exp = oracle.bfs(row, col, src)
    gx.tune_set("bfs.visited_bitmap", vis)
    try:
        a, st = gx.bfs(G, src)
        b, _ = gx.bfs(G, 7)
    finally:
        gx.tune_set("bfs.visited_bitmap", -1)
    np.testing.assert_array_equal(a.cpu().numpy().view(np.uint32), exp)
    np.testing.assert_array_equal(b.cpu().numpy().view(np.uint32), oracle.bfs(row, col, 7))


@pytest.mark.parametrize("hub", [64, 100000])
def test_hub_degree_knob(gx, oracle, rmat16, hub):
    """Results do not depend on the grid-wide hub threshold."""
    G, row, col, w = rmat16
    s = int(np.argmax(np.diff(row)))
    say(f"rmat16 hub_degree={hub}")
    gx.tune_set("frontier.hub_degree", hub)
    try:
        lv, _ = gx.bfs(G, s)
        d, _ = gx.sssp(G, s, 6)
        lab, _ = gx.connected_components(G)
        deg, _, _, _ = gx.kcore(G, 10)
    finally:
        gx.tune_set("frontier.hub_degree", -1)
    np.testing.assert_array_equal(u32(lv), oracle.bfs(row, col, s, threads=8))
    np.testing.assert_array_equal(u32(d), oracle.sssp(row, col, w, s, 6, threads=8))
    np.testing.assert_array_equal(u32(lab), oracle.cc(row, col, threads=8))
    np.testing.assert_array_equal(u32(deg), oracle.kcore(row, col, 10, threads=8))


@pytest.mark.parametrize("merge,defer,refine", [(1, 0, 0), (4096, 0, 0), (256, 1, 0), (256, 0, 1), (1, 0, 2),
                                                 (256, 0, 7), (256, 0, 32)])
def test_sssp_bucket_merge_knob(gx, oracle, rmat16, merge, defer, refine):
    """Distances do not depend on how many buckets a phase may merge
    (sssp.max_merge; 1 = one bucket per phase), on deferring hubs
    (sssp.defer_hubs) or on the distance slices hub steps expand
    (sssp.hub_refine; 1 = one distance per hub step, 32 = all pending)."""
    G, row, col, w = rmat16
    Gg = gx.Graph.grid(256)
    grow, gcol, gw = Gg.to_host()
    say(f"sssp max_merge={merge} defer_hubs={defer} hub_refine={refine}")
    gx.tune_set("sssp.max_merge", merge)
    gx.tune_set("sssp.defer_hubs", defer)
    gx.tune_set("sssp.hub_refine", refine)
    try:
        d, _ = gx.sssp(G, 0, 6)
        dg, stg = gx.sssp(Gg, 0, 6)
        dg2, _ = gx.sssp(Gg, 77, 0)
    finally:
        gx.tune_set("sssp.max_merge", -1)
        gx.tune_set("sssp.defer_hubs", -1)
        gx.tune_set("sssp.hub_refine", -1)
    np.testing.assert_array_equal(u32(d), oracle.sssp(row, col, w, 0, 6, threads=8))
    np.testing.assert_array_equal(u32(dg), oracle.sssp(grow, gcol, gw, 0, 6, threads=8))
    np.testing.assert_array_equal(u32(dg2), oracle.sssp(grow, gcol, gw, 77, 0, threads=8))


@pytest.mark.parametrize("refine", [0, 2])
def test_degree_ordered_copy_matches_oracle(gx, oracle, rmat16, refine):
    """BFS, direction-optimising BFS, SSSP and k-core on the degree-ordered
    copy of the CSR (frontier.relabel = 1; on by default only from 2^24
    vertices) give the oracle's levels, distances and core degrees in the
    caller's vertex ids, and valid parents, with and without the
    distance-ordered hub steps (sssp.hub_refine)."""
    G, row, col, w = rmat16
    srcs = [0, int(np.argmax(np.diff(row))), 12345]
    say(f"rmat16 frontier.relabel=1 hub_refine={refine} sources={srcs}")
    gx.tune_set("frontier.relabel", 1)
    gx.tune_set("sssp.hub_refine", refine)
    try:
        got = [(u32(gx.bfs(G, s)[0]), u32(gx.sssp(G, s, 6)[0])) for s in srcs]
        dos = [tuple(u32(t) for t in gx.bfs_do(G, s, parents=True)[:2]) for s in srcs]
        deg, alive, count, _ = gx.kcore(G, 10)
    finally:
        gx.tune_set("frontier.relabel", -1)
        gx.tune_set("sssp.hub_refine", -1)
    n = len(row) - 1
    src_of = np.repeat(np.arange(n, dtype=np.int64), np.diff(row))
    edge_keys = np.sort(src_of * n + col.astype(np.int64))
    for s, (lv, d), (lv2, par) in zip(srcs, got, dos):
        exp = oracle.bfs(row, col, s, threads=8)
        np.testing.assert_array_equal(lv, exp, err_msg=f"bfs from {s}")
        np.testing.assert_array_equal(lv2, exp, err_msg=f"direction-optimising bfs from {s}")
        np.testing.assert_array_equal(d, oracle.sssp(row, col, w, s, 6, threads=8), err_msg=f"sssp from {s}")
        # parents in the caller's ids: one level up and joined by an edge
        reached = exp != INF
        assert par[s] == s and np.all(par[~reached] == 0xFFFFFFFF)
        v = np.nonzero(reached & (np.arange(n) != s))[0]
        p = par[v].astype(np.int64)
        assert np.all(exp[p] == exp[v] - 1)
        keys = p * n + v
        at = np.searchsorted(edge_keys, keys)
        assert np.all(at < len(edge_keys)) and np.all(edge_keys[np.minimum(at, len(edge_keys) - 1)] == keys)
    exp = oracle.kcore(row, col, 10, threads=8)
    np.testing.assert_array_equal(u32(deg), exp)
    np.testing.assert_array_equal(alive.cpu().numpy(), (exp >= 10).astype(np.uint8))
    assert count == int((exp >= 10).sum())


def _sssp_rows_match(gx, oracle, G, row, col, w, srcs, shift):
    got, st = gx.sssp_batched(G, srcs, shift)
    got = got.cpu().numpy().view(np.uint32)
    assert got.shape == (len(srcs), G.n)
    for i, s in enumerate(srcs):
        np.testing.assert_array_equal(got[i], oracle.sssp(row, col, w, int(s), shift, threads=8),
                                      err_msg=f"source {s} (row {i})")
    return st


def test_sssp_batched_golden(gx, golden):
    for g in golden["undirected"]:
        G = dev(gx, g)
        srcs = [int(s) for s in g["sssp"]]
        for shift in (0, 6):
            say(f"sssp batched {g['name']} shift={shift} sources={srcs}")
            got, _ = gx.sssp_batched(G, srcs, shift)
            got = got.cpu().numpy().view(np.uint32)
            for i, s in enumerate(srcs):
                np.testing.assert_array_equal(got[i], np.asarray(g["sssp"][str(s)], np.uint32), err_msg=g["name"])


def test_sssp_batched_rmat16_grid_directed(gx, oracle, rmat16):
    """150 seeded sources on RMAT-16 (two full groups and a partial one, a
    repeated source), 70 on a 128-grid, 40 on a directed RMAT-12: every row
    bit-exact against the oracle's single-source delta-stepping."""
    G, row, col, w = rmat16
    srcs = gx.pick_sources(row, 150)
    srcs[7] = srcs[3]  # a source twice in one group
    say("sssp batched rmat16 150 sources shift 6")
    st = _sssp_rows_match(gx, oracle, G, row, col, w, srcs, 6)
    assert st.iterations > 0
    say("sssp batched grid128 70 sources shift 6")
    Gg = gx.Graph.grid(128)
    grow, gcol, gw = Gg.to_host()
    _sssp_rows_match(gx, oracle, Gg, grow, gcol, gw, gx.pick_sources(grow, 70), 6)
    say("sssp batched directed rmat12 40 sources shift 4")
    Gd = gx.Graph.rmat(12, 8, symmetric=False, weights="uniform", wlo=1, whi=255)
    drow, dcol, dw = Gd.to_host()
    _sssp_rows_match(gx, oracle, Gd, drow, dcol, dw, np.arange(0, 4000, 100, dtype=np.uint32), 4)
    # no sources: nothing to do
    got, _ = gx.sssp_batched(G, np.zeros(0, np.uint32), 6)
    assert got.shape == (0, G.n)


def test_grid_parity(gx, oracle):
    say("grid 256 parity")
    G = gx.Graph.grid(256)
    row, col, w = G.to_host()
    d, _ = gx.sssp(G, 0, 6)
    np.testing.assert_array_equal(u32(d), oracle.sssp(row, col, w, 0, 6, threads=8))
    lab, _ = gx.connected_components(G)
    np.testing.assert_array_equal(u32(lab), oracle.cc(row, col, threads=8))
    deg, _, _, _ = gx.kcore(G, 3)
    np.testing.assert_array_equal(u32(deg), oracle.kcore(row, col, 3, threads=8))
    lv, _ = gx.bfs(G, 0)
    np.testing.assert_array_equal(u32(lv), oracle.bfs(row, col, 0, threads=8))


# ------------------------------------------------------------- edge cases
def test_edge_cases(gx):
    from galois_amd import _lib
    say("edge cases")
    # one vertex, no edges
    G = gx.Graph.from_csr(np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint32),
                          symmetric=True)
    lv, _ = gx.bfs(G, 0)
    assert u32(lv).tolist() == [0]
    d, _ = gx.sssp(G, 0, 6)
    assert u32(d).tolist() == [0]
    lab, _ = gx.connected_components(G)
    assert u32(lab).tolist() == [0]
    deg, alive, cnt, _ = gx.kcore(G, 1)
    assert cnt == 0
    x, _ = gx.pagerank(G, precision="fp64")
    assert abs(float(x[0]) - 1.0) < 1e-12
    # isolated vertices around a single edge
    row = np.array([0, 0, 1, 2, 2], np.int32)
    col = np.array([2, 1], np.int32)
    G = gx.Graph.from_csr(row, col, np.array([5, 5], np.uint32), symmetric=True)
    assert u32(gx.bfs(G, 1)[0]).tolist() == [INF, 0, 1, INF]
    assert u32(gx.sssp(G, 2, 0)[0]).tolist() == [INF, 5, 0, INF]
    assert u32(gx.connected_components(G)[0]).tolist() == [0, 1, 1, 3]
    # errors: out-of-range source, unweighted SSSP, asymmetric CC
    with pytest.raises(_lib.GxError) as e:
        gx.bfs(G, 9)
    assert e.value.code == _lib.GX_ERR_INVALID
    U = gx.Graph.from_csr(row, col, None, symmetric=False)
    with pytest.raises(_lib.GxError):
        gx.sssp(U, 0)
    with pytest.raises(_lib.GxError):
        gx.connected_components(U)
    with pytest.raises(_lib.GxError):
        gx.kcore(U, 2)


def test_repeat_runs_are_identical(gx, rmat16):
    G, row, col, w = rmat16
    say("repeat determinism")
    a = u32(gx.sssp(G, 3, 6)[0]).copy()
    b = u32(gx.sssp(G, 3, 6)[0]).copy()
    np.testing.assert_array_equal(a, b)
    x1 = gx.pagerank(G, precision="fp32")[0].cpu().numpy().copy()
    x2 = gx.pagerank(G, precision="fp32")[0].cpu().numpy().copy()
    np.testing.assert_array_equal(x1, x2)  # fixed-order reductions: deterministic


# ------------------------------------------------------------- PageRank shapes
def test_pagerank_long_rows_and_empty_runs(gx, oracle):
    """Rows longer than a window (piece sweep), windows with more rows than
    the LDS row table (lane-per-row fallback), dangling vertices."""
    say("pagerank long rows / empty runs")
    rng = np.random.default_rng(7)
    n = 20000
    src = list(range(1, 5000)) + list(rng.integers(5000, 6000, 4000))
    dst = [0] * 4999 + list(rng.integers(5000, 6000, 4000))
    src += [0, 0, 17]
    dst += [19999, 5000, 12000]
    keys = np.unique(np.array(src, np.int64) * n + np.array(dst, np.int64))
    s, d = keys // n, keys % n
    keep = s != d
    s, d = s[keep], d[keep]
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, s + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    col = d.astype(np.int32)
    G = gx.Graph.from_csr(row, col, None, symmetric=False)
    exp, it, _ = oracle.pagerank(row, col, tol=1e-12, max_iter=200)
    x64, st = gx.pagerank(G, tol=1e-12, max_iter=200, precision="fp64")
    assert st.iterations == it
    np.testing.assert_allclose(x64.cpu().numpy(), exp, rtol=1e-11, atol=1e-18)
    x32, _ = gx.pagerank(G, tol=1e-12, max_iter=200, precision="fp32")
    assert np.abs(x32.cpu().numpy().astype(np.float64) - exp).sum() < 1e-6


# ------------------------------------------------------------- BC
def _bc_close(got, exp, rel=1e-6):
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    scale = np.maximum(np.abs(exp), 1.0)
    assert np.all(np.abs(got - exp) <= rel * scale), np.max(np.abs(got - exp) / scale)


def test_bc_golden(gx, golden):
    for g in golden["undirected"]:
        G = dev(gx, g, weights=False)
        n = g["n"]
        for fn in (gx.betweenness_centrality, gx.betweenness_centrality_batched):
            say(f"bc {g['name']} {fn.__name__}")
            bc, _ = fn(G, np.arange(n))
            _bc_close(bc.cpu().numpy(), g["bc_all"])
            bc, _ = fn(G, g["bc_sources"])
            _bc_close(bc.cpu().numpy(), g["bc_subset"])


def test_bc_directed_and_duplicates(gx, golden, oracle):
    for g in golden["directed"]:
        row, col, _ = csr_arrays(g)
        G = dev(gx, g, symmetric=False, weights=False)
        srcs = np.array([0, 3, 3, 7, 11, 0], np.uint32)
        exp = oracle.bc(row, col, srcs)
        for fn in (gx.betweenness_centrality, gx.betweenness_centrality_batched):
            say(f"bc directed {g['name']} {fn.__name__}")
            _bc_close(fn(G, srcs)[0].cpu().numpy(), exp)


def test_bc_rmat16_plain_vs_batched(gx, oracle, rmat16):
    G, row, col, w = rmat16
    srcs = gx.pick_sources(row, 150)
    say("bc rmat16 oracle")
    exp = oracle.bc(row, col, srcs, threads=8)
    say("bc rmat16 plain")
    plain, _ = gx.betweenness_centrality(G, srcs)
    _bc_close(plain.cpu().numpy(), exp)
    say("bc rmat16 batched")
    batched, st = gx.betweenness_centrality_batched(G, srcs)
    _bc_close(batched.cpu().numpy(), exp)


@pytest.mark.parametrize("chunk,grid", [(64, 256), (1024, 16384)])
def test_bc_batched_knobs(gx, oracle, rmat16, chunk, grid):
    """Batched BC is the same at any chunk length (degree-bin boundary) and
    sweep grid: RMAT-16 (150 sources) and a union of 6 uneven graphs."""
    G, row, col, w = rmat16
    srcs = gx.pick_sources(row, 150)
    parts = [gx.Graph.rmat(sc, ef, seed=900 + i, symmetric=True) for i, (sc, ef) in
             enumerate([(9, 8), (11, 16), (7, 4), (10, 32), (8, 8), (12, 8)])]
    U, offs = gx.Graph.union(parts)
    urow, ucol, _ = U.to_host()
    usrc = np.concatenate([gx.pick_sources(P.to_host()[0], 5, seed=i) + offs[i] for i, P in enumerate(parts)])
    say(f"bc batched bc.chunk={chunk} bc.grid={grid}")
    gx.tune_set("bc.chunk", chunk)
    gx.tune_set("bc.grid", grid)
    try:
        got, _ = gx.betweenness_centrality_batched(G, srcs)
        ugot, _ = gx.betweenness_centrality_batched(U, usrc.astype(np.uint32))
    finally:
        gx.tune_set("bc.chunk", -1)
        gx.tune_set("bc.grid", -1)
    _bc_close(got.cpu().numpy(), oracle.bc(row, col, srcs, threads=8))
    _bc_close(ugot.cpu().numpy(), oracle.bc(urow, ucol, usrc, threads=8))


# ------------------------------------------------------------- differentiable PageRank
def _norm_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_weighted_pagerank_golden(gx, golden):
    import torch
    for g in golden["weighted"]:
        say(f"weighted pagerank {g['name']}")
        row = np.asarray(g["row"], np.int32)
        col = np.asarray(g["col"], np.int32)
        w = np.asarray(g["w"], np.float32)
        G = gx.Graph.from_csr(row, col, w, symmetric=False)
        wt = torch.tensor(w, device="cuda")
        t = torch.tensor(np.asarray(g["t"], np.float32), device="cuda")
        c = torch.tensor(np.asarray(g["c"], np.float32), device="cuda")
        x64, _ = gx.pagerank_weighted(G, wt, t, alpha=g["alpha"], tol=1e-15, max_iter=2000, precision="fp64")
        # the golden system was built from the same float32 weights
        np.testing.assert_allclose(x64.cpu().numpy(), g["x"], rtol=1e-6)
        gw, gt, _ = gx.pagerank_adjoint(G, wt, x64, c, alpha=g["alpha"], tol=1e-15, max_iter=2000,
                                        precision="fp64")
        np.testing.assert_allclose(gw.cpu().numpy(), g["grad_w"], rtol=1e-4, atol=1e-7)
        np.testing.assert_allclose(gt.cpu().numpy(), g["grad_t"], rtol=1e-4, atol=1e-7)


def test_adjoint_vs_oracle_and_finite_differences(gx, oracle):
    import torch
    say("adjoint rmat12 lognormal")
    G = gx.Graph.rmat(12, 8, symmetric=False, weights="lognormal")
    row, col, w = G.to_host()
    n, m = G.n, G.m
    rng = np.random.default_rng(5)
    t = rng.random(n).astype(np.float32)
    t /= t.sum()
    c = rng.normal(size=n).astype(np.float32)
    wt, tt, ct = (torch.tensor(a, device="cuda") for a in (w, t, c))
    x_or = oracle.pagerank_weighted(row, col, w, t, alpha=0.85, tol=1e-15, max_iter=3000)
    gw_or, gt_or = oracle.pagerank_adjoint(row, col, w, x_or, c, alpha=0.85, tol=1e-15, max_iter=3000)
    x64, _ = gx.pagerank_weighted(G, wt, tt, tol=1e-15, max_iter=3000, precision="fp64")
    np.testing.assert_allclose(x64.cpu().numpy(), x_or, rtol=1e-9)
    gw64, gt64, _ = gx.pagerank_adjoint(G, wt, x64, ct, tol=1e-15, max_iter=3000, precision="fp64")
    np.testing.assert_allclose(gw64.cpu().numpy(), gw_or, rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(gt64.cpu().numpy(), gt_or, rtol=1e-8, atol=1e-14)
    # FP32 forward + adjoint against the FP64 gradient: 1e-5 (norm-relative)
    x32, st_f = gx.pagerank_weighted(G, wt, tt, tol=1e-9, max_iter=300, precision="fp32")
    gw32, gt32, st_b = gx.pagerank_adjoint(G, wt, x32, ct, tol=1e-9, max_iter=300, precision="fp32")
    assert _norm_rel(gw32.cpu().numpy(), gw_or) < 1e-5
    assert _norm_rel(gt32.cpu().numpy(), gt_or) < 1e-5
    # central finite differences (eps = 1e-4) of the FP64 oracle on sampled edges
    idx = rng.choice(m, 40, replace=False)
    for e in idx:
        wp = w.astype(np.float64).copy()
        wm = wp.copy()
        wp[e] += 1e-4
        wm[e] -= 1e-4
        fd = (np.dot(c, oracle.pagerank_weighted(row, col, wp, t, tol=1e-15, max_iter=3000))
              - np.dot(c, oracle.pagerank_weighted(row, col, wm, t, tol=1e-15, max_iter=3000))) / 2e-4
        assert abs(fd - gw64.cpu().numpy()[e]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9


def test_autograd_function(gx):
    import torch
    say("autograd")
    G = gx.Graph.rmat(10, 8, symmetric=False, weights="lognormal")
    _, _, w = G.to_host()
    rng = np.random.default_rng(9)
    w = torch.tensor(w, device="cuda", requires_grad=True)
    t = torch.full((G.n,), 1.0 / G.n, device="cuda", requires_grad=True)
    c = torch.tensor(rng.normal(size=G.n).astype(np.float32), device="cuda")
    x = gx.differentiable_pagerank(G, w, t, tol=1e-9, max_iter=300)
    loss = (c * x).sum()
    loss.backward()
    gw, gt, _ = gx.pagerank_adjoint(G, w.detach(), x.detach(), c, tol=1e-9, max_iter=300)
    np.testing.assert_array_equal(w.grad.cpu().numpy(), gw.cpu().numpy())
    np.testing.assert_array_equal(t.grad.cpu().numpy(), gt.cpu().numpy())


# ------------------------------------------------------------- uneven graphs in one launch
def test_batched_bc_over_union_of_uneven_graphs(gx, oracle):
    say("union of uneven graphs")
    rng = np.random.default_rng(3)
    parts = [gx.Graph.rmat(int(rng.integers(7, 12)), int(rng.integers(4, 17)), seed=100 + i, symmetric=True)
             for i in range(12)]
    U, offs = gx.Graph.union(parts)
    srcs = []
    exp = np.zeros(U.n)
    for i, P in enumerate(parts):
        row, col, _ = P.to_host()
        s = gx.pick_sources(row, 3, seed=7 + i)
        srcs.extend((s + offs[i]).tolist())
        exp[offs[i]:offs[i + 1]] = oracle.bc(row, col, s)
    bc, st = gx.betweenness_centrality_batched(U, np.array(srcs, np.uint32))
    _bc_close(bc.cpu().numpy(), exp)


def test_batched_bc_scattered_components(gx, oracle):
    """Batched BC sweeps each 64-source group over the id span of its sources'
    components.  Here the union's vertex ids are shuffled, so every span also
    holds vertices of other components (and groups mix graphs): results must
    still equal the oracle's, and a second call (reused level buffers) too."""
    say("batched bc, scattered components")
    rng = np.random.default_rng(5)
    parts = [gx.Graph.rmat(int(rng.integers(7, 11)), int(rng.integers(4, 13)), seed=300 + i, symmetric=True)
             for i in range(10)]
    U, _ = gx.Graph.union(parts)
    del parts
    row, col, _ = U.to_host()
    n = len(row) - 1
    perm = rng.permutation(n)  # old id -> new id
    src_old = np.repeat(np.arange(n), np.diff(row))
    s_new, d_new = perm[src_old], perm[col]
    order = np.lexsort((d_new, s_new))
    s_new, d_new = s_new[order], d_new[order]
    prow = np.zeros(n + 1, np.int64)
    np.add.at(prow, s_new + 1, 1)
    prow = np.cumsum(prow).astype(np.int32)
    pcol = d_new.astype(np.int32)
    G = gx.Graph.from_csr(prow, pcol, None, symmetric=True)
    srcs = gx.pick_sources(prow, 150, seed=11)
    exp = oracle.bc(prow, pcol, srcs)
    for _ in range(2):
        bc, _ = gx.betweenness_centrality_batched(G, srcs)
        _bc_close(bc.cpu().numpy(), exp)
    # a group whose sources sit in one component, after groups that did not
    one = srcs[np.argsort(prow[srcs + 1] - prow[srcs])][:3]
    bc1, _ = gx.betweenness_centrality_batched(G, one)
    _bc_close(bc1.cpu().numpy(), oracle.bc(prow, pcol, one))


def test_autotuner_picks_and_caches(gx, tmp_path):
    from galois_amd import autotune
    say("autotune rmat14")
    G = gx.Graph.rmat(14, 16, symmetric=True)
    cache = tmp_path / "tune.json"
    try:
        c1 = autotune.tune(G, "bfs", cache=cache, source=0)
        assert c1["frontier.hub_degree"] in autotune.CANDIDATES["frontier.hub_degree"]
        assert gx.tune_get("frontier.hub_degree") == c1["frontier.hub_degree"]
        c2 = autotune.tune(G, "pagerank", cache=cache, iterations=2)
        # the hot/cold split is timed only when the merge-path pull runs
        if c2["pr.blocking"] == 0:
            assert c2["pr.hot_sources"] in autotune.CANDIDATES["pr.hot_sources"]
        else:
            assert c2["pr.hot_sources"] == -1 and gx.tune_get("pr.hot_sources") == -1
        c3 = autotune.tune(G, "pagerank", cache=cache, iterations=2, precision="fp64")
        assert c3["pr.blocking"] == -1 and c3["pr.hot_sources"] in autotune.CANDIDATES["pr.hot_sources"]
        row, _, _ = G.to_host()
        c4 = autotune.tune(G, "bc", cache=cache, sources=gx.pick_sources(row, 70))
        assert c4["bc.chunk"] in autotune.CANDIDATES["bc.chunk"] and c4["bc.grid"] in autotune.CANDIDATES["bc.grid"]
        assert gx.tune_get("bc.chunk") == c4["bc.chunk"]
        Gw = gx.Graph.rmat(14, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
        c5 = autotune.tune(Gw, "sssp", cache=cache, source=0)
        for knob in ("frontier.hub_degree", "sssp.hub_refine", "sssp.max_merge"):
            assert c5[knob] in autotune.CANDIDATES[knob] and gx.tune_get(knob) == c5[knob], knob
        autotune.reset()
        assert autotune.tune(G, "bfs", cache=cache, source=0) == c1  # from the cache
        assert len(__import__("json").loads(cache.read_text())) == 5  # FP32 and FP64 keyed apart
    finally:
        autotune.reset()
