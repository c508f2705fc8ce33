"""GPU parity at the full sizes of the benchmark workloads, through the C ABI,
against the oracle (the CPU restatement of the lonestar apps) on the same
seeded graphs:

  1. BFS from the max-degree vertex on symmetric RMAT-18 (levels bit-exact;
     top-down and direction-optimising).
  2. SSSP (delta 64 = shift 6) from vertex 0, min-label CC and k-core (k=10)
     on symmetric RMAT-22 with weights 1..255 and on the 2048 x 2048 road-like
     grid (5% of lattice edges dropped, 1% shortcuts, weights 1..1000), all
     bit-exact.
  3. PageRank FP32 (default blocked sweep) and FP64 on directed RMAT-26
     against the oracle's FP64 pull: L1 and L-infinity stated.
  4. Batched BC: 256 seeded sources on symmetric RMAT-22, and all 256 sources
     (4 per graph) of 64 uneven RMAT graphs (scales 12-18, edge factors 8-32
     from the seed) in one union, to 1e-6 relative against the FP64 oracle.
  5. Adjoint PageRank on directed weighted RMAT-22 (lognormal weights): FP64
     and FP32 gradients against the oracle's FP64 gradient, and central
     finite differences (eps 1e-4) of the GPU FP64 forward on 1000 sampled
     edges.
  6. BFS, direction-optimising BFS, SSSP, CC and k-core on symmetric RMAT-26
     (2.1e9 edges), bit-exact.

Sizes are the workloads' own; the oracle runs on up to 16 host threads."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 27491095  # lonestar/liblonestar/include/Lonestar/Utils.h:31
INF = 2147483646
THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 4))))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def _bc_close(got, exp, rel=1e-6):
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    scale = np.maximum(np.abs(exp), 1.0)
    err = np.abs(got - exp) / scale
    assert np.all(err <= rel), float(err.max())
    return float(err.max())


# ----------------------------------------------------------------- config 1
def test_config1_bfs_rmat18(gx, oracle):
    say("config 1: bfs symmetric rmat18 from the max-degree vertex")
    G = gx.Graph.rmat(18, 16, symmetric=True)
    row, col, _ = G.to_host()
    assert G.n == 1 << 18 and G.m > 7_000_000
    src = int(np.argmax(np.diff(row)))
    exp = oracle.bfs(row, col, src, THREADS)
    np.testing.assert_array_equal(u32(gx.bfs(G, src)[0]), exp)
    np.testing.assert_array_equal(u32(gx.bfs_do(G, src)[0]), exp)


# ----------------------------------------------------------------- config 2
def _graph2(gx, name):
    if name == "rmat22":
        return gx.Graph.rmat(22, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    return gx.Graph.grid(2048, 50000, 10000, SEED, 1, 1000)


@pytest.mark.parametrize("name", ["rmat22", "grid2048"])
def test_config2_sssp_cc_kcore(gx, oracle, name):
    say(f"config 2: sssp/cc/kcore {name}")
    G = _graph2(gx, name)
    row, col, w = G.to_host()
    assert G.n == 1 << 22
    say(f"  n={G.n} m={G.m}: sssp shift 6 from 0")
    np.testing.assert_array_equal(u32(gx.sssp(G, 0, 6)[0]), oracle.sssp(row, col, w, 0, 6, THREADS))
    say("  cc")
    np.testing.assert_array_equal(u32(gx.connected_components(G)[0]), oracle.cc(row, col, THREADS))
    say("  kcore 10")
    deg, alive, cnt, _ = gx.kcore(G, 10)
    exp = oracle.kcore(row, col, 10, THREADS)
    np.testing.assert_array_equal(u32(deg), exp)
    assert cnt == int((exp >= 10).sum())
    np.testing.assert_array_equal(alive.cpu().numpy().astype(bool), exp >= 10)


# ----------------------------------------------------------------- config 3
def test_config3_pagerank_rmat26(gx, oracle):
    """FP32 (exact fixed-point row sums) and FP64 PageRank on directed RMAT-26
    against the oracle's FP64 pull to the same tolerance (1e-9 L1 residual,
    at most 100 iterations).  Stated: L1 and L-infinity of each against the
    oracle.  FP32 must be within 1e-6 in L1, FP64 within 1e-10."""
    say("config 3: pagerank directed rmat26 fp32/fp64 vs oracle fp64")
    G = gx.Graph.rmat(26, 16, symmetric=False, seed=SEED)
    x32, st32 = gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp32")
    x64, st64 = gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp64")
    a = x32.cpu().numpy().astype(np.float64)
    b = x64.cpu().numpy()
    del x32, x64
    row, col, _ = G.to_host()
    del G
    say(f"  oracle fp64 pull on n={len(row) - 1} m={len(col)}")
    ref, it, res = oracle.pagerank(row, col, threads=THREADS)
    l1_32, li_32 = float(np.abs(a - ref).sum()), float(np.abs(a - ref).max())
    l1_64, li_64 = float(np.abs(b - ref).sum()), float(np.abs(b - ref).max())
    say(f"  fp32: L1 {l1_32:.3e} Linf {li_32:.3e} ({st32.iterations} it); fp64: L1 {l1_64:.3e} "
        f"Linf {li_64:.3e} ({st64.iterations} it); oracle {it} it")
    assert st64.iterations == it
    assert l1_32 < 1e-6 and li_32 < 1e-9
    assert l1_64 < 1e-10 and li_64 < 1e-12
    assert abs(a.sum() - 1.0) < 1e-5 and abs(b.sum() - 1.0) < 1e-12


# ----------------------------------------------------------------- config 4
def test_config4a_bc_256_sources_rmat22(gx, oracle):
    say("config 4a: batched bc, 256 seeded sources, symmetric rmat22")
    G = gx.Graph.rmat(22, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    row, col, _ = G.to_host()
    srcs = gx.pick_sources(row, 256)
    assert len(np.unique(srcs)) == 256
    got, _ = gx.betweenness_centrality_batched(G, srcs)
    got = got.cpu().numpy()
    say("  oracle bc")
    exp = oracle.bc(row, col, srcs, THREADS)
    say(f"  max rel err {_bc_close(got, exp):.3e}")


def uneven_graphs(gx, ngraphs=64):
    """64 RMAT graphs with scale in [12, 18] and edge factor in [8, 32] drawn
    from the seed, as one disjoint union, with 4 seeded sources per graph."""
    rng = np.random.default_rng(SEED)
    specs = [(int(rng.integers(12, 19)), int(rng.integers(8, 33))) for _ in range(ngraphs)]
    parts = [gx.Graph.rmat(sc, ef, seed=SEED + 1000 + i, symmetric=True) for i, (sc, ef) in enumerate(specs)]
    U, offs = gx.Graph.union(parts)
    srcs = []
    for i, P in enumerate(parts):
        prow, _, _ = P.to_host()
        srcs.extend((gx.pick_sources(prow, 4, seed=SEED + i) + offs[i]).tolist())
    return U, offs, np.array(srcs, np.uint32), specs


def test_config4b_bc_uneven64(gx, oracle):
    say("config 4b: batched bc over 64 uneven graphs, 4 sources each")
    U, offs, srcs, specs = uneven_graphs(gx)
    assert len(srcs) == 256 and len(offs) == 65
    got, _ = gx.betweenness_centrality_batched(U, srcs)
    got = got.cpu().numpy()
    row, col, _ = U.to_host()
    say(f"  oracle bc on the union n={U.n} m={U.m}")
    exp = oracle.bc(row, col, srcs, THREADS)
    say(f"  max rel err {_bc_close(got, exp):.3e}")


# ----------------------------------------------------------------- config 5
@pytest.fixture(scope="module")
def adjoint_case(gx):
    """Directed RMAT-22 with lognormal(0,1) FP32 weights (row-normalised by the
    kernels), uniform teleport, c ~ N(0,1) from the seed; GPU FP64 forward and
    FP64/FP32 adjoints."""
    import torch
    say("config 5: weighted directed rmat22")
    G = gx.Graph.rmat(22, 16, symmetric=False, weights="lognormal", seed=SEED + 5)
    n = G.n
    gen = torch.Generator(device="cuda").manual_seed(SEED)
    c = torch.randn(n, generator=gen, device="cuda", dtype=torch.float32)
    t = torch.full((n,), 1.0 / n, device="cuda", dtype=torch.float32)
    x64, _ = gx.pagerank_weighted(G, None, t, tol=1e-14, max_iter=400, precision="fp64")
    x32, _ = gx.pagerank_weighted(G, None, t, tol=1e-10, max_iter=400, precision="fp32")
    gw64, gt64, _ = gx.pagerank_adjoint(G, None, x64, c, tol=1e-14, max_iter=400, precision="fp64")
    gw32, gt32, _ = gx.pagerank_adjoint(G, None, x32, c, tol=1e-10, max_iter=400, precision="fp32")
    return {"G": G, "c": c, "t": t, "x64": x64, "gw64": gw64.cpu().numpy(), "gt64": gt64.cpu().numpy(),
            "gw32": gw32.cpu().numpy(), "gt32": gt32.cpu().numpy()}


def test_config5_adjoint_rmat22_vs_oracle(adjoint_case, oracle):
    """The oracle's FP64 forward and adjoint (threaded) on the same graph: the
    GPU FP64 forward within 1e-10 in L1 (both solved to an L1 residual of
    1e-14; they differ in the FP64 summation order), the FP64 gradients
    within 1e-9 (norm-relative), the FP32 gradients within 1e-5
    (norm-relative)."""
    A = adjoint_case
    G = A["G"]
    row, col, w = G.to_host()
    w64 = w.view(np.float32).astype(np.float64)
    c = A["c"].cpu().numpy().astype(np.float64)
    t = A["t"].cpu().numpy().astype(np.float64)
    say(f"  oracle forward fp64 n={G.n} m={G.m}")
    x_ref = oracle.pagerank_weighted(row, col, w64, t, tol=1e-14, max_iter=400, threads=THREADS)
    x64 = A["x64"].cpu().numpy()
    l1 = float(np.abs(x64 - x_ref).sum())
    say(f"  forward fp64 L1 vs oracle {l1:.3e}")
    assert l1 < 1e-10
    say("  oracle adjoint fp64")
    gw_ref, gt_ref = oracle.pagerank_adjoint(row, col, w64, x_ref, c, tol=1e-14, max_iter=400, threads=THREADS)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    r64w, r64t = rel(A["gw64"], gw_ref), rel(A["gt64"], gt_ref)
    r32w, r32t = rel(A["gw32"], gw_ref), rel(A["gt32"], gt_ref)
    say(f"  grad_w rel err fp64 {r64w:.3e} fp32 {r32w:.3e}; grad_t fp64 {r64t:.3e} fp32 {r32t:.3e}")
    assert r64w < 1e-9 and r64t < 1e-9
    assert r32w < 1e-5 and r32t < 1e-5


FD_EDGES = 1000
FD_PARTS = 4
FD_ITERS = 120  # fixed iteration count for both perturbed solves (0.85^120 = 3e-9)


@pytest.mark.parametrize("part", range(FD_PARTS))
def test_config5_adjoint_rmat22_finite_differences(gx, adjoint_case, part):
    """Central differences (eps = 1e-4 on one FP32 edge weight) of L = c.x with
    the GPU FP64 forward, on 1000 edges sampled with probability proportional
    to |dL/dw| (seeded), 250 per part.  Both perturbed solves run the same
    fixed number of iterations from the same start, so the difference
    quotient converges to the derivative like 0.85^k; the truncation error of
    the central difference is O(eps^2).  Each must be within 1e-4 relative
    (plus 1e-13 absolute) of the GPU FP64 adjoint."""
    import torch
    A = adjoint_case
    G, c, t, g = A["G"], A["c"], A["t"], A["gw64"]
    rng = np.random.default_rng(SEED + 55)
    p = np.abs(g) / np.abs(g).sum()
    edges = rng.choice(len(g), size=FD_EDGES, replace=False, p=p)
    mine = edges[part::FD_PARTS]
    _, _, w = G.to_host()
    wt = torch.from_numpy(w.view(np.float32).copy()).cuda()
    c64 = c.double()
    worst = 0.0
    for k, e in enumerate(mine):
        e = int(e)
        vals, ws = [], []
        for sgn in (1.0, -1.0):
            w2 = wt.clone()
            w2[e] += sgn * 1e-4
            ws.append(float(w2[e]))  # the FP32 weight actually used
            xe, _ = gx.pagerank_weighted(G, w2, t, tol=0.0, max_iter=FD_ITERS, precision="fp64")
            vals.append(float((xe * c64).sum()))
        fd = (vals[0] - vals[1]) / (ws[0] - ws[1])
        err = abs(fd - g[e])
        worst = max(worst, err / abs(g[e]))
        assert err <= 1e-4 * abs(g[e]) + 1e-13, (e, fd, g[e])
        if k % 50 == 0:
            say(f"  part {part}: {k}/{len(mine)} edges, worst rel err {worst:.2e}")
    say(f"  part {part}: {len(mine)} edges, worst rel err {worst:.2e}")


# ----------------------------------------------------------------- RMAT-26
def test_rmat26_integer_kernels_bit_exact(gx, oracle):
    """BFS and direction-optimising BFS from the max-degree vertex, SSSP
    (shift 6) from vertex 0, CC and k-core (k=10) on symmetric RMAT-26 with
    weights 1..255 (about 2.1e9 directed edges), bit-exact against the
    oracle."""
    say("rmat26 symmetric weighted: bfs, bfs_do, sssp, cc, kcore vs oracle")
    G = gx.Graph.rmat(26, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    row, col, w = G.to_host()
    src = int(np.argmax(np.diff(row)))
    say(f"  n={G.n} m={G.m} src={src}: bfs")
    lv = u32(gx.bfs(G, src)[0])
    exp = oracle.bfs(row, col, src, THREADS)
    np.testing.assert_array_equal(lv, exp)
    np.testing.assert_array_equal(u32(gx.bfs_do(G, src)[0]), exp)
    del lv, exp
    say("  cc")
    np.testing.assert_array_equal(u32(gx.connected_components(G)[0]), oracle.cc(row, col, THREADS))
    say("  kcore 10")
    np.testing.assert_array_equal(u32(gx.kcore(G, 10)[0]), oracle.kcore(row, col, 10, THREADS))
    say("  sssp shift 6 from 0")
    np.testing.assert_array_equal(u32(gx.sssp(G, 0, 6)[0]), oracle.sssp(row, col, w, 0, 6, THREADS))
