#!/usr/bin/env python3
"""Benchmark: lonestar analytics on one MI355X.

Headline (`value`): PageRank, FP32, on a directed RMAT scale-26 graph
(67,108,864 vertices, edge factor 16, a/b/c/d = .57/.19/.19/.05, seeded,
self-loops and duplicate edges removed), damping 0.85, uniform teleport,
dangling mass redistributed, stop at L1 residual < 1e-9 or 100 iterations,
through the default propagation-blocked sweep (exact fixed-point row sums).
One step = one full PageRank solve; value = GTEPS = edges x iterations / time.

Also reported, one line of JSON in total:
  roofline      the dominant kernel (one sweep: phase 1 + phase 2 + sliced-bin
                finalisation) against the 8 TB/s HBM peak, with the pull
                iteration's algorithmic bytes 4m + 16n + 4 per sweep, timed
                with HIP events on the library's stream during the timed
                steps; beside them the bytes of the sweep actually run as the
                library states them (blocked: 2m + 10g + 20n + 12 per job,
                g = (source, destination bin) groups) and measured HBM
                traffic from the committed rocprofv3 summary;
  cpu_baseline  the CPU restatement (oracle, FP64 pull) timed on this host's
                cores on a bounded sample: 3 PageRank iterations on the same
                graph, median of 5 after a warm-up;
  kernels       GTEPS and %HBM of every kernel on the other workloads (BFS on
                symmetric RMAT-18; SSSP/CC/k-core on symmetric RMAT-22 and the
                2048x2048 road-like grid; batched vs one-source-per-pass BC,
                on RMAT-22 and on 64 uneven graphs with 4 and with 64
                sources per graph, with the autotuner's batched-BC choices;
                batched vs one-source SSSP; adjoint vs forward PageRank; the
                D-IrGL personality loops; BFS/SSSP/CC/k-core on symmetric
                RMAT-26), each beside the oracle's time on this host's cores
                (cpu_ms: median after a warm-up) and whether the outputs are
                equal, plus the PageRank error against the CPU FP64 solve.
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
SEED = 27491095    # lonestar/liblonestar/include/Lonestar/Utils.h:31


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def sync():
    import torch
    torch.cuda.synchronize()


def frontier_traffic() -> dict:
    """Measured HBM bytes per run of each frontier algorithm on symmetric
    RMAT-26 from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE passes of `--frontier-pass --frontier-only <algorithm>`,
    calibrated against a copy kernel), or {} when it is absent."""
    f = ROOT / "profiles" / "frontier_rmat26_traffic.json"
    try:
        return {k: v["bytes_per_run"] for k, v in json.loads(f.read_text())["per_algorithm"].items()}
    except (OSError, ValueError, KeyError, TypeError):
        return {}


def measured_traffic(scale: int, sweep: str):
    """HBM bytes per PageRank sweep from the committed rocprofv3 PMC summary
    of that sweep (FETCH_SIZE and WRITE_SIZE passes, calibrated against a copy
    kernel), or None when no summary for this graph and sweep is present."""
    f = ROOT / "profiles" / f"pagerank_{sweep}_rmat26_traffic.json"
    if scale != 26 or not f.exists():
        return None
    try:
        return json.loads(f.read_text())["traffic_bytes_per_sweep"]
    except (ValueError, KeyError):
        return None


# ------------------------------------------------------------ CPU baseline
CPU_THREADS = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def oracle():
    """The CPU restatement (tests/conftest.py's wrapper of oracle/liboracle.so):
    the checker and the timed CPU baseline, never the product."""
    sys.path.insert(0, str(ROOT / "tests"))
    from conftest import ORACLE_SO, Oracle
    return Oracle(ORACLE_SO)


def cpu_time(fn, reps: int):
    """Median wall time (ms) of `reps` runs after one warm-up, and the last
    result (the reference's Timer_0 region: the algorithm only)."""
    out = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts)), out


def cpu_row(ms: float, work: int, reps: int, equal=None) -> dict:
    r = {"cpu_ms": round(ms, 3), "cpu_GTEPS": round(work / ms / 1e6, 4), "cpu_threads": CPU_THREADS,
         "cpu_runs": f"median of {reps} after a warm-up"}
    if equal is not None:
        r["equal_oracle"] = bool(equal)
    return r


def cpu_baseline_pagerank(G, iters: int, reps: int = 5) -> dict:
    """Times the oracle's FP64 pull PageRank (CSC given) for `iters` iterations
    on the host cores, median of `reps` after a warm-up."""
    so = ROOT / "oracle" / "liboracle.so"
    lib = C.CDLL(str(so))
    fn = lib.oracle_pagerank_fp64_csc
    fn.restype = C.c_int
    fn.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_uint32,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    row = G.row_host()
    trow, tcol, _ = G.transpose_host()
    outdeg = np.diff(row).astype(np.uint32)
    x = np.empty(G.n, np.float64)
    it = C.c_uint32()
    res = C.c_double()

    def run():
        assert fn(G.n, outdeg.ctypes.data, trow.ctypes.data, tcol.ctypes.data, 0.85, 0.0, iters, x.ctypes.data,
                  C.byref(it), C.byref(res), CPU_THREADS) == 0
    ms, _ = cpu_time(run, reps)
    return {"value": round(G.m * it.value / ms / 1e6, 4), "unit": "GTEPS", "cores": CPU_THREADS, "kind": "port",
            "sample": f"{it.value} FP64 pull PageRank iterations on the same RMAT-26 graph ({G.m} edges), "
                      f"median of {reps} after a warm-up: {ms / 1e3:.2f} s, threads={CPU_THREADS}"}


# ------------------------------------------------------------ kernel table
def pct(bytes_, ms):
    return round(100.0 * bytes_ / (ms * 1e-3) / HBM_PEAK, 2)


def kernel_table(gx, quick: bool, tuned: dict) -> dict:
    out = {}
    O = oracle()
    u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    # Config 1: BFS on symmetric RMAT-18 from the max-degree vertex.
    G = gx.Graph.rmat(18, 16, symmetric=True)
    row, col, _ = G.to_host()
    n, m = G.n, G.m
    src = int(np.argmax(np.diff(row)))
    cms, exp = cpu_time(lambda: O.bfs(row, col, src, CPU_THREADS), 5)
    for name, fn in (("bfs_rmat18", lambda: gx.bfs(G, src)), ("bfs_do_rmat18", lambda: gx.bfs_do(G, src))):
        fn()
        best = min((fn() for _ in range(3)), key=lambda r: r[-1].elapsed_ms)
        st = best[-1]
        lv = u32(best[0])
        n_r = int((lv != gx.GX_DIST_INFINITY).sum())
        m_r = st.edges_visited
        byt = 4 * (n + 1) + 4 * m_r + 4 * n + 4 * n_r
        out[name] = {"ms": round(st.elapsed_ms, 4), "GTEPS": round(m_r / st.elapsed_ms / 1e6, 3),
                     "pct_hbm": pct(byt, st.elapsed_ms), "levels": st.iterations, "n": n, "m": m,
                     **cpu_row(cms, m_r, 5, np.array_equal(lv, exp))}
    del G
    # Config 2A / 2B: SSSP (delta 64 = shift 6), CC, k-core (k=10).
    graphs = [("rmat22", lambda: gx.Graph.rmat(22, 16, symmetric=True, weights="uniform", wlo=1, whi=255)),
              ("grid2048", lambda: gx.Graph.grid(2048, 50000, 10000, SEED, 1, 1000))]
    if quick:
        graphs = [("rmat18w", lambda: gx.Graph.rmat(18, 16, symmetric=True, weights="uniform", wlo=1, whi=255)),
                  ("grid256", lambda: gx.Graph.grid(256, 50000, 10000, SEED, 1, 1000))]
    for gname, make in graphs:
        G = make()
        row, col, wts = G.to_host()
        n, m = G.n, G.m
        deg = np.diff(row).astype(np.int64)
        runs = {"sssp": lambda: gx.sssp(G, 0, 6), "cc": lambda: gx.connected_components(G),
                "kcore10": lambda: gx.kcore(G, 10)}
        cpu = {"sssp": lambda: O.sssp(row, col, wts, 0, 6, CPU_THREADS), "cc": lambda: O.cc(row, col, CPU_THREADS),
               "kcore10": lambda: O.kcore(row, col, 10, CPU_THREADS)}
        for kname, fn in runs.items():
            fn()
            res = min((fn() for _ in range(3)), key=lambda r: r[-1].elapsed_ms)
            st = res[-1]
            extra = {}
            if kname == "sssp":
                d = u32(res[0])
                m_r = int(deg[d != gx.GX_DIST_INFINITY].sum())
                byt = 8 * m_r + 12 * n
                extra = {"edges_relaxed": st.edges_visited, "relaxed_per_edge": round(st.edges_visited / max(m, 1), 3)}
            elif kname == "cc":
                byt = 4 * m + 12 * n
            else:
                alive = res[1].cpu().numpy()
                m_dead = int(deg[alive == 0].sum())
                byt = 4 * (n + 1) + 4 * m_dead + 8 * n
            cms, exp = cpu_time(cpu[kname], 5)
            out[f"{kname}_{gname}"] = {"ms": round(st.elapsed_ms, 4), "GTEPS": round(m / st.elapsed_ms / 1e6, 3),
                                       "pct_hbm": pct(byt, st.elapsed_ms), "rounds": st.iterations, "n": n,
                                       "m": m, **extra, **cpu_row(cms, m, 5, np.array_equal(u32(res[0]), exp))}
        if gname.startswith("rmat"):
            # Residual push PageRank (pagerank_push.cpp, tolerance 1e-6 as the
            # D-IrGL row) against the oracle's residual pull, the same
            # unnormalised fixed point.
            gx.pagerank_push(G, tol=1e-6)
            pp = min((gx.pagerank_push(G, tol=1e-6) for _ in range(3)), key=lambda r: r[-1].elapsed_ms)
            ref, _ = O.pagerank_residual(row, col, tol=1e-6, threads=CPU_THREADS)
            got = pp[0].cpu().numpy()
            out[f"pagerank_push_{gname}"] = {
                "ms": round(pp[-1].elapsed_ms, 3), "rounds": pp[-1].iterations,
                "max_rel_diff_vs_oracle_residual_pull": float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)))}
            del pp, got, ref
        # Config 4a: BC on the symmetric RMAT-22, 256 seeded sources.
        if gname.startswith("rmat"):
            nsrc = 64 if quick else 256
            srcs = gx.pick_sources(row, nsrc)
            # Batched SSSP: 64 seeded sources at once vs one source per call,
            # 8 of the rows checked against the oracle.
            bs, stbs = gx.sssp_batched(G, srcs[:64], 6)
            bs, stbs = gx.sssp_batched(G, srcs[:64], 6)
            one = [gx.sssp(G, int(x), 6)[1].elapsed_ms for x in srcs[:8]]
            bsh = bs.cpu().numpy().view(np.uint32)
            eq = all(np.array_equal(bsh[i], O.sssp(row, col, wts, int(srcs[i]), 6, CPU_THREADS)) for i in range(8))
            cms, _ = cpu_time(lambda: O.sssp(row, col, wts, int(srcs[0]), 6, CPU_THREADS), 5)
            out[f"sssp_batched_{gname}"] = {
                "sources": 64, "ms": round(stbs.elapsed_ms, 3), "rounds": stbs.iterations,
                "plain_ms_per_source": round(float(np.mean(one)), 3),
                "speedup_batched_vs_plain": round(float(np.mean(one)) * 64 / stbs.elapsed_ms, 2),
                "rows_equal_oracle_8_sources": bool(eq),
                "cpu_ms_per_source": round(cms, 3), "cpu_threads": CPU_THREADS}
            del bs, bsh
            from galois_amd import autotune
            t0 = time.perf_counter()
            tuned[f"bc_{gname}"] = {"choice": autotune.tune(G, "bc", cache=None, sources=srcs),
                                    "seconds": round(time.perf_counter() - t0, 1)}
            gx.betweenness_centrality_batched(G, srcs[:64])
            bat, stb = gx.betweenness_centrality_batched(G, srcs)
            nplain = 16 if quick else 32
            gx.betweenness_centrality(G, srcs[:4])
            pl, stp = gx.betweenness_centrality(G, srcs[:nplain])
            ref_sub, _ = gx.betweenness_centrality_batched(G, srcs[:nplain])
            a, b = pl.cpu().numpy(), ref_sub.cpu().numpy()
            rel = float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1.0)))
            per_src_plain = stp.elapsed_ms / nplain
            byt = (8 * m + 40 * n) * nsrc
            cms, _ = cpu_time(lambda: O.bc(row, col, srcs[:16], CPU_THREADS), 3)
            out[f"bc_batched_{gname}"] = {"sources": nsrc, "ms": round(stb.elapsed_ms, 3),
                                          "TEPS_G": round(nsrc * m / stb.elapsed_ms / 1e6, 3),
                                          "pct_hbm_vs_unbatched_bytes": pct(byt, stb.elapsed_ms),
                                          "knobs": {k: gx.tune_get(k) for k in ("bc.chunk", "bc.grid")},
                                          "cpu_ms_per_source": round(cms / 16, 3), "cpu_threads": CPU_THREADS,
                                          "cpu_runs": "16 sources, median of 3 after a warm-up"}
            from galois_amd import autotune
            autotune.reset()
            out[f"bc_plain_{gname}"] = {"sources_timed": nplain, "ms_per_source": round(per_src_plain, 3),
                                        "TEPS_G": round(m / per_src_plain / 1e6, 3)}
            out[f"bc_speedup_batched_vs_plain_{gname}"] = round(per_src_plain * nsrc / stb.elapsed_ms, 2)
            bc_or = O.bc(row, col, srcs[:nplain], CPU_THREADS)
            out[f"bc_max_rel_err_vs_oracle_{nplain}_sources_{gname}"] = {
                "plain": float(np.max(np.abs(a - bc_or) / np.maximum(np.abs(bc_or), 1.0))),
                "batched": float(np.max(np.abs(b - bc_or) / np.maximum(np.abs(bc_or), 1.0)))}
            out.update(personality_rows(gx, row, col, wts, gname))
        del G
    # Config 4b: one batch of 64 uneven graphs (RMAT scales 12-18, edge
    # factors 8-32 from the seed), 4 seeded sources per graph, batched BC over
    # their disjoint union vs one source per pass; checked on a subset
    # against the oracle (FP64 Brandes).
    rng = np.random.default_rng(SEED)
    ngraphs = 16 if quick else 64
    specs = [(int(rng.integers(12, 19)), int(rng.integers(8, 33))) for _ in range(ngraphs)]
    if quick:
        specs = [(min(sc, 14), ef) for sc, ef in specs]
    parts = [gx.Graph.rmat(sc, ef, seed=SEED + 1000 + i, symmetric=True) for i, (sc, ef) in enumerate(specs)]
    U, offs = gx.Graph.union(parts)
    srcs = []
    for i, P in enumerate(parts):
        prow, _, _ = P.to_host()
        srcs.extend((gx.pick_sources(prow, 4, seed=SEED + i) + offs[i]).tolist())
    del parts
    srcs = np.array(srcs, np.uint32)
    from galois_amd import autotune
    t0 = time.perf_counter()
    tuned["bc_uneven64"] = {"choice": autotune.tune(U, "bc", cache=None, sources=srcs),
                            "seconds": round(time.perf_counter() - t0, 1)}
    gx.betweenness_centrality_batched(U, srcs[:64])
    bat, stb = gx.betweenness_centrality_batched(U, srcs)
    nplain = 16
    pl, stp = gx.betweenness_centrality(U, srcs[:nplain])
    per_src_plain = stp.elapsed_ms / nplain
    sub, _ = gx.betweenness_centrality_batched(U, srcs[:8])
    urow, ucol, _ = U.to_host()
    import ctypes as _C
    lib = _C.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    exp = np.empty(U.n, np.float64)
    lib.oracle_bc.argtypes = [_C.c_int64, _C.c_void_p, _C.c_void_p, _C.c_void_p, _C.c_uint32, _C.c_void_p, _C.c_int]
    assert lib.oracle_bc(U.n, urow.ctypes.data, ucol.ctypes.data, srcs[:8].ctypes.data, 8, exp.ctypes.data,
                         int(os.environ.get("OMP_NUM_THREADS", "8"))) == 0
    got = sub.cpu().numpy()
    out["bc_batched_uneven64"] = {
        "graphs": ngraphs, "sum_n": int(U.n), "sum_m": int(U.m), "sources": int(len(srcs)),
        "ms": round(stb.elapsed_ms, 3), "TEPS_G": round(len(srcs) * U.m / ngraphs / stb.elapsed_ms / 1e6, 3),
        "plain_ms_per_source": round(per_src_plain, 3),
        "speedup_batched_vs_plain": round(per_src_plain * len(srcs) / stb.elapsed_ms, 2),
        "max_rel_err_vs_oracle_8_sources": float(np.max(np.abs(got - exp) / np.maximum(np.abs(exp), 1.0))),
        "knobs": {k: gx.tune_get(k) for k in ("bc.chunk", "bc.grid")}}
    cms, _ = cpu_time(lambda: O.bc(urow, ucol, srcs, CPU_THREADS), 3)
    out["bc_batched_uneven64"].update({"cpu_ms": round(cms, 3), "cpu_threads": CPU_THREADS,
                                       "cpu_runs": "all 256 sources, median of 3 after a warm-up"})
    # The same 64 graphs with 64 seeded sources each (4096 sources): every
    # 64-source batch then lies in one graph, so all its lanes work at every
    # vertex it reaches (with 4 sources per graph a batch spans 16 graphs and
    # at most 4 of its 64 lanes are live at a vertex).  One source per pass is
    # timed on 16 of them; 8 are checked against the oracle.
    srcs64 = []
    for i in range(ngraphs):
        lo, hi = int(offs[i]), int(offs[i + 1])
        srcs64.extend((gx.pick_sources(urow[lo:hi + 1] - urow[lo], 64, seed=SEED + 7000 + i) + lo).tolist())
    srcs64 = np.array(srcs64, np.uint32)
    gx.betweenness_centrality_batched(U, srcs64[:64])
    bat64, stb64 = gx.betweenness_centrality_batched(U, srcs64)
    pl64, stp64 = gx.betweenness_centrality(U, srcs64[::len(srcs64) // 16][:16])
    per_src_plain64 = stp64.elapsed_ms / 16
    sub64, _ = gx.betweenness_centrality_batched(U, srcs64[:8])
    exp64 = np.empty(U.n, np.float64)
    assert lib.oracle_bc(U.n, urow.ctypes.data, ucol.ctypes.data, srcs64[:8].ctypes.data, 8, exp64.ctypes.data,
                         int(os.environ.get("OMP_NUM_THREADS", "8"))) == 0
    got64 = sub64.cpu().numpy()
    out["bc_batched_uneven64_64_per_graph"] = {
        "graphs": ngraphs, "sources": int(len(srcs64)), "ms": round(stb64.elapsed_ms, 3),
        "TEPS_G": round(len(srcs64) * U.m / ngraphs / stb64.elapsed_ms / 1e6, 3),
        "plain_ms_per_source": round(per_src_plain64, 3),
        "speedup_batched_vs_plain": round(per_src_plain64 * len(srcs64) / stb64.elapsed_ms, 2),
        "max_rel_err_vs_oracle_8_sources": float(np.max(np.abs(got64 - exp64) / np.maximum(np.abs(exp64), 1.0)))}
    autotune.reset()
    del U, bat, pl, sub, bat64, pl64, sub64
    # Config 5: differentiable PageRank on directed weighted RMAT-22 (FP32
    # lognormal weights, row-normalised), loss sum(c * x), c ~ N(0,1).
    scale5 = 16 if quick else 22
    import torch
    G = gx.Graph.rmat(scale5, 16, symmetric=False, weights="lognormal", seed=SEED + 5)
    n, m = G.n, G.m
    gen = torch.Generator(device="cuda").manual_seed(SEED)
    c = torch.randn(n, generator=gen, device="cuda", dtype=torch.float32)
    t = torch.full((n,), 1.0 / n, device="cuda", dtype=torch.float32)
    res = {}
    for prec in ("fp32", "fp64"):
        gx.pagerank_weighted(G, None, t, tol=1e-10, max_iter=200, precision=prec)
        x, stf = gx.pagerank_weighted(G, None, t, tol=1e-10, max_iter=200, precision=prec)
        gw, gt, sta = gx.pagerank_adjoint(G, None, x, c, tol=1e-10, max_iter=200, precision=prec)
        res[prec] = (x, gw, gt, stf, sta)
    _, gw32, gt32, stf32, sta32 = res["fp32"]
    x64, gw64, gt64, _, _ = res["fp64"]
    a64, b64 = gw64.cpu().numpy(), gt64.cpu().numpy()
    rel_w = float(np.linalg.norm(gw32.cpu().numpy() - a64) / np.linalg.norm(a64))
    rel_t = float(np.linalg.norm(gt32.cpu().numpy() - b64) / np.linalg.norm(b64))
    # Central differences (FP64, eps 1e-4) on the 16 edges of largest gradient.
    _, _, w_host = G.to_host()
    top = np.argsort(-np.abs(a64))[:16]
    wt = torch.from_numpy(w_host.astype(np.float32)).cuda()
    c64 = c.double()
    fd = []
    for e in top:
        vals, ws = [], []
        for sgn in (1.0, -1.0):
            w2 = wt.clone()
            w2[int(e)] += sgn * 1e-4
            ws.append(float(w2[int(e)]))  # the float32 weight actually used
            xe, _ = gx.pagerank_weighted(G, w2, t, tol=1e-14, max_iter=400, precision="fp64")
            vals.append(float((xe * c64).sum()))
        fd.append((vals[0] - vals[1]) / (ws[0] - ws[1]))
    fd = np.array(fd)
    out["pagerank_adjoint_rmat22"] = {
        "n": n, "m": m, "forward_fp32_ms": round(stf32.elapsed_ms, 3), "adjoint_fp32_ms": round(sta32.elapsed_ms, 3),
        "adjoint_over_forward": round(sta32.elapsed_ms / stf32.elapsed_ms, 3),
        "forward_iterations": stf32.iterations, "adjoint_iterations": sta32.iterations,
        "grad_w_rel_err_fp32_vs_fp64": rel_w, "grad_t_rel_err_fp32_vs_fp64": rel_t,
        "fd_max_rel_err_16_edges": float(np.max(np.abs(fd - a64[top]) / np.abs(a64[top])))}
    del G
    if not quick:
        out.update(rmat26_table(gx, O))
    return out


def personality_rows(gx, row, col, wts, gname) -> dict:
    """The D-IrGL host loops of one host through the personality partition
    (gx_part_*): BFS push from the max-degree vertex, SSSP push from vertex 0,
    CC push, PageRank push, BC level (16 sources) and k-core push (k=10),
    timed around the whole loop (every
    operator returns its reductions to the host, as in the reference); the
    partition upload is not timed.  Results are checked against the oracle:
    BFS levels, SSSP distances, CC labels and k-core membership bit-exact,
    PageRank push against the oracle's FP32 residual pull (the same
    unnormalised fixed point, tolerance 1e-6), BC against the
    FP64 Brandes oracle."""
    from galois_amd import personality as pers
    out = {}
    n, m = len(row) - 1, len(col)
    P = pers.Partition(row, col, weights=wts)
    src = int(np.argmax(np.diff(row)))

    def timed(fn):
        best, res = None, None
        for _ in range(3):
            t0 = time.perf_counter()
            res = fn()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best, res

    O = oracle()
    dt, (dist, rounds, work) = timed(lambda: pers.dirgl_sssp(P, 0))
    d = O.sssp(row, col, wts, 0, 6, CPU_THREADS)  # shortest distances do not depend on the bucket width
    same = bool(np.array_equal(np.where(d == gx.GX_DIST_INFINITY, pers.INFINITY, d), dist))
    out[f"dirgl_sssp_push_{gname}"] = {"ms": round(dt * 1e3, 3), "GTEPS": round(m / dt / 1e9, 3), "rounds": rounds,
                                       "work_edges": work, "dist_equal_oracle": same}
    dt, (label, comps, rounds) = timed(lambda: pers.dirgl_cc(P))
    lab = O.cc(row, col, CPU_THREADS)
    out[f"dirgl_cc_push_{gname}"] = {"ms": round(dt * 1e3, 3), "GTEPS": round(m / dt / 1e9, 3), "rounds": rounds,
                                     "components": comps, "labels_equal_oracle": bool(np.array_equal(lab, label))}
    # PageRank push at the reference app's default tolerance 1e-6
    # (pagerank_push.cpp:48-50); compared with the library's residual push on
    # the nodes with edges (the others keep 0 in the D-IrGL loop).
    dt, (value, rounds) = timed(lambda: pers.dirgl_pagerank(P, 1e-6))
    ref, _ = O.pagerank_residual(row, col, tol=1e-6, threads=CPU_THREADS)  # same unnormalised fixed point
    we = P.range(pers.RANGE_WITH_EDGES)[1]
    rel = np.abs(value[:we] - ref[:we]) / np.maximum(np.abs(ref[:we]), 1.0)
    out[f"dirgl_pagerank_push_{gname}"] = {"ms": round(dt * 1e3, 3), "rounds": rounds,
                                           "max_rel_diff_vs_oracle_residual_pull": float(rel.max())}
    # BC level over 16 seeded sources against the FP64 Brandes oracle.
    srcs = [int(x) for x in gx.pick_sources(row, 16)]
    dt, bcv = timed(lambda: pers.dirgl_bc(P, srcs))
    ref = O.bc(row, col, np.asarray(srcs, np.uint32), CPU_THREADS)
    out[f"dirgl_bc_level_{gname}"] = {"sources": len(srcs), "ms": round(dt * 1e3, 3),
                                      "ms_per_source": round(dt * 1e3 / len(srcs), 3),
                                      "max_rel_diff_vs_oracle": float(np.max(np.abs(bcv - ref)
                                                                             / np.maximum(np.abs(ref), 1.0)))}
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        dist, rounds, work = pers.dirgl_bfs(P, src)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    lv = O.bfs(row, col, src, CPU_THREADS)
    same = bool(np.array_equal(np.where(lv == gx.GX_DIST_INFINITY, pers.INFINITY, lv), dist))
    out[f"dirgl_bfs_push_{gname}"] = {"ms": round(best * 1e3, 3), "GTEPS": round(m / best / 1e9, 3), "rounds": rounds,
                                      "work_items": work, "levels_equal_oracle": same}
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        flag, alive, rounds = pers.dirgl_kcore(P, 10)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    # Nodes past the last node with edges are not visited by KCoreStep1 and
    # keep their flag (the reference's D-IrGL semantics).
    we = P.range(pers.RANGE_WITH_EDGES)[1]
    ka = O.kcore(row, col, 10, CPU_THREADS) >= 10
    same = (bool(np.array_equal(flag[:we].astype(bool), ka[:we])) and bool(flag[we:].all())
            and alive == int(ka[:we].sum()) + (n - we))
    out[f"dirgl_kcore10_push_{gname}"] = {"ms": round(best * 1e3, 3), "GTEPS": round(m / best / 1e9, 3),
                                          "rounds": rounds, "alive": alive, "membership_equal_oracle": same}
    P.free()
    return out


def rmat26_table(gx, O) -> dict:
    """BFS (top-down and direction-optimising, from the max-degree vertex),
    SSSP (delta 64 from vertex 0), CC and k-core (k=10) on the symmetrised
    RMAT-26 graph with uniform weights 1..255 (2.1e9 directed edges): times,
    GTEPS (m / time, the Graph500 convention for the direction-optimising BFS)
    and the algorithmic-byte share of HBM peak, beside the oracle's time on
    this host's cores (median of 3 after a warm-up) and whether the outputs
    are bit-equal."""
    out = {}
    G = gx.Graph.rmat(26, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    row, col, wts = G.to_host()
    n, m = G.n, G.m
    deg = np.diff(row).astype(np.int64)
    src = int(np.argmax(deg))
    cpu = {"bfs": lambda: O.bfs(row, col, src, CPU_THREADS), "sssp": lambda: O.sssp(row, col, wts, 0, 6, CPU_THREADS),
           "cc": lambda: O.cc(row, col, CPU_THREADS), "kcore10": lambda: O.kcore(row, col, 10, CPU_THREADS)}
    ref = {}
    traffic = frontier_traffic()
    runs = {"bfs": lambda: gx.bfs(G, src), "bfs_do": lambda: gx.bfs_do(G, src), "sssp": lambda: gx.sssp(G, 0, 6),
            "cc": lambda: gx.connected_components(G), "kcore10": lambda: gx.kcore(G, 10)}
    for kname, fn in runs.items():
        fn()
        res = min((fn() for _ in range(2)), key=lambda r: r[-1].elapsed_ms)
        st = res[-1]
        extra = {}
        if kname.startswith("bfs"):
            lv = res[0].cpu().numpy().view(np.uint32)
            reached = lv != gx.GX_DIST_INFINITY
            byt = 4 * (n + 1) + 4 * int(deg[reached].sum()) + 4 * n + 4 * int(reached.sum())
        elif kname == "sssp":
            d = res[0].cpu().numpy().view(np.uint32)
            byt = 8 * int(deg[d != gx.GX_DIST_INFINITY].sum()) + 12 * n
            extra = {"edges_relaxed": st.edges_visited, "relaxed_per_edge": round(st.edges_visited / max(m, 1), 3)}
        elif kname == "cc":
            byt = 4 * m + 12 * n
        else:
            alive = res[1].cpu().numpy()
            byt = 4 * (n + 1) + 4 * int(deg[alive == 0].sum()) + 8 * n
        base = "bfs" if kname.startswith("bfs") else kname
        if base not in ref:
            ref[base] = cpu_time(cpu[base], 3)
        cms, exp = ref[base]
        tkey = "kcore" if kname == "kcore10" else kname
        if tkey in traffic:  # measured bytes of the same kernels (profiles/frontier_rmat26_traffic.json)
            extra.update({"pct_hbm_basis": "algorithmic bytes (work-equivalent for DO-BFS and CC)",
                          "measured_hbm_GB": round(traffic[tkey] / 1e9, 2),
                          "pct_hbm_measured": pct(traffic[tkey], st.elapsed_ms)})
        out[f"{kname}_rmat26"] = {"ms": round(st.elapsed_ms, 4), "GTEPS": round(m / st.elapsed_ms / 1e6, 3),
                                  "pct_hbm": pct(byt, st.elapsed_ms), "rounds": st.iterations, "n": n, "m": m,
                                  **extra, **cpu_row(cms, m, 3, np.array_equal(res[0].cpu().numpy().view(np.uint32),
                                                                               exp))}
    del G, row, col, wts
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26, help="RMAT scale of the headline graph")
    ap.add_argument("--quick", action="store_true", help="small graphs for the kernel table")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel table")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set a tuning knob before running (gx_tune_set), e.g. pr.hot_sources=0")
    ap.add_argument("--autotune", action="store_true",
                    help="run galois_amd.autotune on the headline graph before the timed steps")
    ap.add_argument("--frontier-pass", action="store_true",
                    help="workload for rocprofv3 kernel traces and counters of the frontier kernels: k-core "
                         "(k=10), SSSP (delta 64), top-down and direction-optimising BFS and CC on symmetric "
                         "weighted RMAT at --scale, each twice, then a 1 GiB device copy (the calibration "
                         "reference for FETCH_SIZE/WRITE_SIZE); prints no JSON")
    ap.add_argument("--frontier-only", default="",
                    help="comma-separated subset of the frontier pass (kcore,sssp,bfs,bfs_do,cc), so that counters "
                         "of kernels shared by several algorithms can be attributed to one")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="workload for rocprofv3 --pmc runs: 3 pull sweeps on the headline graph, then a 1 GiB "
                         "device copy (the calibration reference for FETCH_SIZE/WRITE_SIZE); prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: there is no multi-GPU path (the kernels shard by "
              "vertex range, a design that is not built); use --gpus 1", file=sys.stderr)
        return 2

    import torch
    import galois_amd as gx
    from galois_amd import _lib
    _lib.check(_lib.lib().gx_init(0))
    torch.cuda.set_device(0)
    for kv in args.tune:
        key, val = kv.split("=", 1)
        gx.tune_set(key, int(val))

    if args.frontier_pass:
        G = gx.Graph.rmat(args.scale, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
        row = G.row_host()
        src = int(np.argmax(np.diff(row)))
        del row
        only = set(args.frontier_only.split(",")) if args.frontier_only else None
        for name, fn in (("kcore", lambda: gx.kcore(G, 10)), ("sssp", lambda: gx.sssp(G, 0, 6)),
                         ("bfs", lambda: gx.bfs(G, src)), ("bfs_do", lambda: gx.bfs_do(G, src)),
                         ("cc", lambda: gx.connected_components(G))):
            if only is not None and name not in only:
                continue
            fn()
            st = fn()[-1]
            log(f"frontier pass: {name} {st.elapsed_ms:.2f} ms, {st.iterations} rounds, "
                f"{st.edges_visited} edges visited")
        del G
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        sync()
        log("frontier pass: 1 GiB copy")
        return 0

    log(f"generating directed RMAT-{args.scale}")
    G = gx.Graph.rmat(args.scale, 16, symmetric=False, seed=SEED)
    n, m = G.n, G.m
    log(f"n={n} m={m}")

    def step():
        return gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp32")

    if args.pmc_pass:
        gx.pagerank(G, alpha=0.85, tol=0.0, max_iter=1, precision="fp32")  # sweep tables built
        _, st = gx.pagerank(G, alpha=0.85, tol=0.0, max_iter=3, precision="fp32")
        del G
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        sync()
        log(f"pmc pass: {st.iterations} sweeps, then a 1 GiB copy")
        return 0

    tuned = {}
    if args.autotune:
        from galois_amd import autotune
        t0 = time.perf_counter()
        tuned = {"pagerank": autotune.tune(G, "pagerank", cache=None, iterations=3)}
        log(f"autotune: {tuned} in {time.perf_counter() - t0:.1f} s")
    for _ in range(args.warmup):
        step()
    sync()
    # The timed steps run with the library's HIP-event timers on (one event
    # pair per pull sweep, on the stream the sweeps are launched on; the
    # production schedule, 8 iterations queued per host check), so the
    # roofline below comes from the same region as `value`.
    gx.set_profiling(True)
    t0 = time.perf_counter()
    iters = sweeps = 0
    sweep_ms_total = 0.0
    for _ in range(args.steps):
        x32, st = step()
        iters += st.iterations
        sweeps += st.main_kernel_launches
        sweep_ms_total += st.main_kernel_ms
        design_bytes = st.main_kernel_bytes
    sync()
    dt = time.perf_counter() - t0
    gx.set_profiling(False)
    value = m * iters / dt / 1e9
    ms_per_step = dt * 1e3 / args.steps
    log(f"pagerank fp32: {ms_per_step:.1f} ms/step, {iters / args.steps:.0f} iterations, {value:.2f} GTEPS")

    # Roofline of the dominant kernel: one FP32 sweep, bracketed by HIP events
    # per iteration on the graph's stream.  Algorithmic bytes per sweep are
    # the pull iteration's minimum for element size 4: 4m source ids + 4(n+1)
    # row offsets + 4n ranks gathered (each once) + 4n ranks written + 4n
    # out-degrees = 4m + 16n + 4.  The bytes of the sweep actually run, as the
    # library states them (gx_stats.main_kernel_bytes), are reported beside:
    #  blocked (default): phase 1 reads a 2-byte source id per group, x and
    #    the out-degrees (8n), writes one 4-byte contribution per group; phase
    #    2 reads a 2-byte destination id per edge, the contributions, its job
    #    table (12 bytes per job) and x_old (4n), writes x' (4n).
    #  pull: 4m column ids + 4(n+1) row offsets + 4n y (read once) + 4n x_old +
    #    4n out-degrees + 8n (x', y' written).
    algo_bytes = 4 * m + 4 * (n + 1) + 12 * n
    blocked = gx.tune_get("pr.blocking") != 0
    sweep = "blocked" if blocked else "pull"
    sweep_ms = sweep_ms_total / max(sweeps, 1)
    if blocked:
        kname = ("propagation-blocked sweep: pb::gather_kernel<float> + pb::accum_kernel<PbFin> "
                 "+ pb::multi_kernel<PbFin> (sliced bins)")
    else:
        kname = ("pull sweep: rs::sweep_kernel<float,HOT,PrSrc,PrFin> + rs::span_finalize_kernel")
    achieved = algo_bytes / (sweep_ms * 1e-3) / 1e9
    design_rate = design_bytes / (sweep_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / (HBM_PEAK / 1e9), 4), "traffic": measured_traffic(args.scale, sweep),
                "kernel": kname, "bytes_per_launch": algo_bytes, "avg_launch_ms": round(sweep_ms, 4),
                "launches": sweeps, "sweep_bytes_per_launch": design_bytes,
                "sweep_bytes_frac": round(design_rate / (HBM_PEAK / 1e9), 4)}

    # FP64 run and the FP32-vs-FP64 error.
    x64, st64 = gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp64")
    a = x32.cpu().numpy().astype(np.float64)
    b = x64.cpu().numpy()
    pr_err = {"l1_fp32_vs_fp64": float(np.abs(a - b).sum()), "linf_fp32_vs_fp64": float(np.abs(a - b).max()),
              "iterations_fp32": st.iterations, "iterations_fp64": st64.iterations,
              "fp64_ms": round(st64.elapsed_ms, 2), "fp64_GTEPS": round(m * st64.iterations / st64.elapsed_ms / 1e6, 3)}
    log(f"fp32 vs fp64: L1 {pr_err['l1_fp32_vs_fp64']:.3e}")

    cpu = None
    layout = gx.pagerank_layout_info(G)
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline_pagerank(G, iters=3)
        # One full FP64 solve on the CPU (tol 1e-9, at most 100 iterations):
        # its time and the GPU runs' error against it.
        row, col, _ = G.to_host()
        t0 = time.perf_counter()
        ref, it_cpu, _ = oracle().pagerank(row, col, threads=CPU_THREADS)
        dt = time.perf_counter() - t0
        del row, col
        pr_err.update({"l1_fp32_vs_cpu_fp64": float(np.abs(a - ref).sum()),
                       "linf_fp32_vs_cpu_fp64": float(np.abs(a - ref).max()),
                       "l1_fp64_vs_cpu_fp64": float(np.abs(b - ref).sum()),
                       "linf_fp64_vs_cpu_fp64": float(np.abs(b - ref).max()), "iterations_cpu": it_cpu,
                       "cpu_full_solve_ms": round(dt * 1e3, 1), "cpu_GTEPS": round(m * it_cpu / dt / 1e9, 4),
                       "cpu_threads": CPU_THREADS, "cpu_runs": "one full solve"})
        del ref
    del G, x32, x64
    torch.cuda.empty_cache()
    kernels = {} if args.no_kernels else kernel_table(gx, args.quick, tuned)
    kernels["pagerank_rmat26"] = pr_err

    out = {
        "metric": "GTEPS & %HBM peak per kernel, RMAT-26 on 1 GPU; batched/adjoint vs plain",
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"pagerank_fp32_rmat{args.scale}_directed", "n": n, "m": m,
                   "edge_factor": 16, "alpha": 0.85, "tol": 1e-9, "max_iter": 100, "seed": SEED,
                   "sweep": "propagation_blocked" if blocked else "merge_path_pull",
                   "knobs": {k: {"set": gx.tune_get(k), "effective": gx.tune_effective(k, n, m)}
                             for k in ("pr.blocking", "pr.hot_sources", "pr.task_edges", "frontier.hub_degree",
                                       "bfs.visited_bitmap", "sssp.max_merge", "sssp.defer_hubs", "sssp.hub_refine", "bc.chunk",
                                       "bc.grid")},
                   "layout": layout,
                   "autotuned": tuned},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
