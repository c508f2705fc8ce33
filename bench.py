#!/usr/bin/env python3
"""Benchmark: lonestar analytics on one MI355X.

Headline (`value`): PageRank pull, FP32 with Kahan-compensated row sums, on a
directed RMAT scale-26 graph (67,108,864 vertices, edge factor 16,
a/b/c/d = .57/.19/.19/.05, seeded, self-loops and duplicate edges removed;
CSR + transposed CSC in HBM), damping 0.85, uniform teleport, dangling mass
redistributed, stop at L1 residual < 1e-9 or 100 iterations.  One step = one
full PageRank solve; value = GTEPS = edges x iterations / time.

Also reported, one line of JSON in total:
  roofline      the dominant kernel (the pull sweep) against the 8 TB/s HBM
                peak, algorithmic bytes 4m + 24n + 4 per sweep, timed with HIP
                events on the library's stream during the timed steps;
  cpu_baseline  the CPU restatement (oracle, FP64 pull) timed on this host's
                cores on a bounded sample (a few PageRank iterations on the
                same graph);
  kernels       GTEPS and %HBM of every kernel on the other workloads (BFS on
                symmetric RMAT-18; SSSP/CC/k-core on symmetric RMAT-22 and the
                2048x2048 road-like grid; batched vs one-source-per-pass BC;
                adjoint vs forward PageRank), plus the FP32-vs-FP64 PageRank
                error.
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


def measured_traffic(scale: int):
    """HBM bytes per pull sweep from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE passes, calibrated against a copy kernel), or
    None when no summary for this graph is present."""
    f = ROOT / "profiles" / "pagerank_pull_rmat26_traffic.json"
    if scale != 26 or not f.exists():
        return None
    try:
        return json.loads(f.read_text())["traffic_bytes_per_sweep"]
    except (ValueError, KeyError):
        return None


# ------------------------------------------------------------ CPU baseline
def cpu_baseline_pagerank(G, iters: int) -> dict:
    """Times the oracle's FP64 pull PageRank (CSC given) for `iters` iterations
    on the host cores.  The oracle is the checker/baseline, never the product."""
    so = ROOT / "oracle" / "liboracle.so"
    lib = C.CDLL(str(so))
    fn = lib.oracle_pagerank_fp64_csc
    fn.restype = C.c_int
    fn.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_uint32,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    row, _, _ = G.to_host()
    trow, tcol, _ = G.transpose_host()
    outdeg = np.diff(row).astype(np.uint32)
    x = np.empty(G.n, np.float64)
    it = C.c_uint32()
    res = C.c_double()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    rc = fn(G.n, outdeg.ctypes.data, trow.ctypes.data, tcol.ctypes.data, 0.85, 0.0, iters, x.ctypes.data,
            C.byref(it), C.byref(res), cores)
    dt = time.perf_counter() - t0
    assert rc == 0
    return {"value": round(G.m * it.value / dt / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
            "sample": f"{it.value} FP64 pull PageRank iterations on the same RMAT-26 graph "
                      f"({G.m} edges), {dt:.2f} s, threads={cores}"}


# ------------------------------------------------------------ kernel table
def pct(bytes_, ms):
    return round(100.0 * bytes_ / (ms * 1e-3) / HBM_PEAK, 2)


def kernel_table(gx, quick: bool) -> dict:
    out = {}
    u32 = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    # Config 1: BFS on symmetric RMAT-18 from the max-degree vertex.
    G = gx.Graph.rmat(18, 16, symmetric=True)
    row, col, _ = G.to_host()
    n, m = G.n, G.m
    src = int(np.argmax(np.diff(row)))
    for name, fn in (("bfs_rmat18", lambda: gx.bfs(G, src)), ("bfs_do_rmat18", lambda: gx.bfs_do(G, src))):
        fn()
        best = min((fn() for _ in range(3)), key=lambda r: r[-1].elapsed_ms)
        st = best[-1]
        lv = u32(best[0])
        n_r = int((lv != gx.GX_DIST_INFINITY).sum())
        m_r = st.edges_visited
        byt = 4 * (n + 1) + 4 * m_r + 4 * n + 4 * n_r
        out[name] = {"ms": round(st.elapsed_ms, 4), "GTEPS": round(m_r / st.elapsed_ms / 1e6, 3),
                     "pct_hbm": pct(byt, st.elapsed_ms), "levels": st.iterations, "n": n, "m": m}
    del G
    # Config 2A / 2B: SSSP (delta 64 = shift 6), CC, k-core (k=10).
    graphs = [("rmat22", lambda: gx.Graph.rmat(22, 16, symmetric=True, weights="uniform", wlo=1, whi=255)),
              ("grid2048", lambda: gx.Graph.grid(2048, 50000, 10000, SEED, 1, 1000))]
    if quick:
        graphs = [("rmat18w", lambda: gx.Graph.rmat(18, 16, symmetric=True, weights="uniform", wlo=1, whi=255)),
                  ("grid256", lambda: gx.Graph.grid(256, 50000, 10000, SEED, 1, 1000))]
    for gname, make in graphs:
        G = make()
        row, col, _ = G.to_host()
        n, m = G.n, G.m
        deg = np.diff(row).astype(np.int64)
        runs = {"sssp": lambda: gx.sssp(G, 0, 6), "cc": lambda: gx.connected_components(G),
                "kcore10": lambda: gx.kcore(G, 10)}
        for kname, fn in runs.items():
            fn()
            res = min((fn() for _ in range(3)), key=lambda r: r[-1].elapsed_ms)
            st = res[-1]
            if kname == "sssp":
                d = u32(res[0])
                m_r = int(deg[d != gx.GX_DIST_INFINITY].sum())
                byt = 8 * m_r + 12 * n
            elif kname == "cc":
                byt = 4 * m + 12 * n
            else:
                alive = res[1].cpu().numpy()
                m_dead = int(deg[alive == 0].sum())
                byt = 4 * (n + 1) + 4 * m_dead + 8 * n
            out[f"{kname}_{gname}"] = {"ms": round(st.elapsed_ms, 4), "GTEPS": round(m / st.elapsed_ms / 1e6, 3),
                                       "pct_hbm": pct(byt, st.elapsed_ms), "rounds": st.iterations, "n": n,
                                       "m": m}
        # Config 4a: BC on the symmetric RMAT-22, 256 seeded sources.
        if gname.startswith("rmat"):
            nsrc = 64 if quick else 256
            srcs = gx.pick_sources(row, nsrc)
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
            out[f"bc_batched_{gname}"] = {"sources": nsrc, "ms": round(stb.elapsed_ms, 3),
                                          "TEPS_G": round(nsrc * m / stb.elapsed_ms / 1e6, 3),
                                          "pct_hbm_vs_unbatched_bytes": pct(byt, stb.elapsed_ms)}
            out[f"bc_plain_{gname}"] = {"sources_timed": nplain, "ms_per_source": round(per_src_plain, 3),
                                        "TEPS_G": round(m / per_src_plain / 1e6, 3)}
            out[f"bc_speedup_batched_vs_plain_{gname}"] = round(per_src_plain * nsrc / stb.elapsed_ms, 2)
            out[f"bc_max_rel_diff_plain_vs_batched_{gname}"] = rel
        del G
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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the per-kernel sharding design is described in DESIGN.md); use --gpus 1", file=sys.stderr)
        return 2

    import torch
    import galois_amd as gx
    from galois_amd import _lib
    _lib.check(_lib.lib().gx_init(0))
    torch.cuda.set_device(0)

    log(f"generating directed RMAT-{args.scale}")
    G = gx.Graph.rmat(args.scale, 16, symmetric=False, seed=SEED)
    n, m = G.n, G.m
    log(f"n={n} m={m}")

    def step():
        return gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp32")

    for _ in range(args.warmup):
        step()
    sync()
    # The timed steps run with the library's HIP-event timers on (one event
    # pair per pull sweep, on the stream the sweeps are launched on), so the
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
    sync()
    dt = time.perf_counter() - t0
    gx.set_profiling(False)
    value = m * iters / dt / 1e9
    ms_per_step = dt * 1e3 / args.steps
    log(f"pagerank fp32: {ms_per_step:.1f} ms/step, {iters / args.steps:.0f} iterations, {value:.2f} GTEPS")

    # Roofline of the dominant kernel, the pull sweep (hot-source pass + cold
    # pass + spanning-row completion, bracketed together per iteration).
    # Algorithmic bytes of one sweep: 4m column ids + 4(n+1) row offsets +
    # 4n y (read once) + 4n x_old + 4n out-degrees + 8n (x', y' written).
    sweep_ms = sweep_ms_total / max(sweeps, 1)
    algo_bytes = 4 * m + 4 * (n + 1) + 20 * n
    achieved = algo_bytes / (sweep_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / (HBM_PEAK / 1e9), 4), "traffic": measured_traffic(args.scale),
                "kernel": "pull sweep: rs::sweep_kernel<float,PrSrc,PartFin> + rs::sweep_kernel<float,PrSrc,PrFin>"
                          " + rs::span_finalize_kernel (x2)",
                "bytes_per_launch": algo_bytes, "avg_launch_ms": round(sweep_ms, 4), "launches": sweeps}

    # FP64 run and the FP32-vs-FP64 error.
    x64, st64 = gx.pagerank(G, alpha=0.85, tol=1e-9, max_iter=100, precision="fp64")
    a = x32.cpu().numpy().astype(np.float64)
    b = x64.cpu().numpy()
    pr_err = {"l1_fp32_vs_fp64": float(np.abs(a - b).sum()), "linf_fp32_vs_fp64": float(np.abs(a - b).max()),
              "iterations_fp32": st.iterations, "iterations_fp64": st64.iterations,
              "fp64_ms": round(st64.elapsed_ms, 2), "fp64_GTEPS": round(m * st64.iterations / st64.elapsed_ms / 1e6, 3)}
    log(f"fp32 vs fp64: L1 {pr_err['l1_fp32_vs_fp64']:.3e}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline_pagerank(G, iters=3)
    del G, x32, x64
    torch.cuda.empty_cache()
    kernels = {} if args.no_kernels else kernel_table(gx, args.quick)
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
        "config": {"workload": f"pagerank_pull_fp32_kahan_rmat{args.scale}_directed", "n": n, "m": m,
                   "edge_factor": 16, "alpha": 0.85, "tol": 1e-9, "max_iter": 100, "seed": SEED},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
