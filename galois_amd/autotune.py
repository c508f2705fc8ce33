"""Autotuner for the library's tuning knobs (gx_tune_set).

For one graph and one algorithm it times every candidate value of each knob
with the library's HIP-event timers (gx_stats.main_kernel_ms, the dominant
kernel's device time; whole-call device time where an entry point times
several kernels), keeps the fastest, and leaves it set.  Choices are cached
per (device, algorithm, workload signature) in a JSON file so a later run on
a similar graph starts from the tuned values without re-timing.

The signature is coarse on purpose: log2 of n, of the average degree and of
the maximum degree, and whether the graph is symmetric — the quantities the
knobs respond to (hub count, gathered working set).

Knobs and candidates:
  frontier.hub_degree  degree above which a frontier vertex is expanded by the
                       whole grid (BFS, SSSP, CC, k-core, BC)
  pr.blocking          FP32 PageRank: propagation-blocked sweep (1) or merge-path pull (0);
                       2 also sends FP64 through the blocked sweep (not a
                       candidate: it changes FP64 results in the last bit)
  pr.hot_sources       sources of the first PageRank pull pass (0 = one pass)
  sssp.max_merge       most delta buckets one SSSP phase may cover
  sssp.hub_refine      SSSP hub steps: 0 = hubs expanded in the step that
                       admits them, r > 0 = deferred and expanded 2^(r-1)
                       distances per hub step
  bc.chunk             batched BC: edges per work item, the degree-bin boundary
                       between whole-adjacency items and split ones
  bc.grid              batched BC: workgroups of each level sweep
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import numpy as np

CANDIDATES: dict[str, list[int]] = {
    "frontier.hub_degree": [64, 128, 256, 512, 1024, 4096],
    "pr.blocking": [1, 0],
    "pr.hot_sources": [0, 1 << 17, 1 << 18, 1 << 19, 1 << 20, 1 << 21],
    "sssp.max_merge": [1, 16, 64, 256, 1024],
    "sssp.hub_refine": [0, 1, 2, 3],
    "bc.chunk": [64, 128, 256, 512, 1024],
    "bc.grid": [2048, 4096, 8192, 16384],
}

# Knobs each algorithm responds to.
KNOBS: dict[str, list[str]] = {
    "bfs": ["frontier.hub_degree"],
    "sssp": ["frontier.hub_degree", "sssp.hub_refine", "sssp.max_merge"],
    "cc": ["frontier.hub_degree"],
    "kcore": ["frontier.hub_degree"],
    "bc": ["bc.chunk", "bc.grid"],
    "bc_plain": ["frontier.hub_degree"],
    "pagerank": ["pr.blocking", "pr.hot_sources"],
}

DEFAULT_CACHE = Path(os.environ.get("XDG_CACHE_HOME", Path.home() / ".cache")) / "galois_amd" / "tune.json"


def signature(gx, G, algo: str, args: dict | None = None) -> str:
    """Coarse workload key: algorithm, device, log2 n, log2 avg/max degree,
    symmetry, and the arguments that pick a different kernel (precision)."""
    import torch
    row, _, _ = G.to_host() if G.m < (1 << 24) else (None, None, None)
    if row is not None:
        deg = np.diff(row)
        dmax = int(deg.max()) if len(deg) else 0
    else:
        dmax = -1  # large graphs: degree extremes are not downloaded
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    sym = bool(G.flags & 1)
    a = args or {}
    variant = a.get("precision", "fp32") if algo == "pagerank" else ""
    if algo == "bc":  # a batch's group count shapes the sweeps
        variant = f"groups={(len(a.get('sources', [0])) + 63) // 64}"
    return json.dumps([algo, dev, round(math.log2(max(G.n, 1))), round(math.log2(max(G.m / max(G.n, 1), 1))),
                       -1 if dmax < 0 else round(math.log2(max(dmax, 1))), sym, variant])


def _applies(knob: str, algo: str, args: dict, choice: dict) -> bool:
    """Whether the kernel that runs responds to `knob`, given the choices made
    so far.  PageRank: pr.blocking candidates (1, 0) differ only for FP32 (FP64
    takes the merge-path pull under both); pr.hot_sources splits the
    merge-path pull only, so it is not timed while the blocked sweep runs."""
    if algo != "pagerank":
        return True
    fp64 = args.get("precision", "fp32") == "fp64"
    if knob == "pr.blocking":
        return not fp64
    if knob == "pr.hot_sources":
        return fp64 or choice.get("pr.blocking", 1) == 0
    return True


def _runner(gx, G, algo: str, args: dict):
    """A callable that runs the algorithm once and returns its device time (ms)."""
    if algo == "bfs":
        src = args.get("source", 0)
        return lambda: gx.bfs(G, src)[1].main_kernel_ms
    if algo == "sssp":
        src, shift = args.get("source", 0), args.get("shift", 6)
        return lambda: gx.sssp(G, src, shift)[1].elapsed_ms
    if algo == "cc":
        return lambda: gx.connected_components(G)[1].elapsed_ms
    if algo == "kcore":
        k = args.get("k", 10)
        return lambda: gx.kcore(G, k)[3].elapsed_ms
    if algo == "bc":
        srcs = args.get("sources", [0])
        return lambda: gx.betweenness_centrality_batched(G, srcs)[1].elapsed_ms
    if algo == "bc_plain":
        srcs = args.get("sources", [0])
        return lambda: gx.betweenness_centrality(G, srcs)[1].elapsed_ms
    if algo == "pagerank":
        prec = args.get("precision", "fp32")
        iters = args.get("iterations", 4)

        def run():
            st = gx.pagerank(G, tol=0.0, max_iter=iters, precision=prec)[1]
            return st.main_kernel_ms / max(st.main_kernel_launches, 1)
        return run
    raise ValueError(f"unknown algorithm {algo!r}")


def tune(G, algo: str, repeats: int = 2, cache: Path | None = DEFAULT_CACHE, retune: bool = False,
         **args) -> dict:
    """Tune the knobs of `algo` on graph G; returns {knob: value} and leaves the
    values set.  Timing uses profiling mode (HIP events per dominant launch)."""
    import galois_amd as gx
    knobs = KNOBS[algo]
    key = signature(gx, G, algo, args)
    table = {}
    if cache is not None and cache.exists():
        try:
            table = json.loads(cache.read_text())
        except ValueError:
            table = {}
    if not retune and key in table:
        for k, v in table[key]["choice"].items():
            gx.tune_set(k, v)
        return dict(table[key]["choice"])
    run = _runner(gx, G, algo, args)
    choice, timings = {}, {}
    gx.set_profiling(True)
    try:
        run()  # warm-up: plans, transposes, captured graphs
        for knob in knobs:
            if not _applies(knob, algo, args, choice):
                gx.tune_set(knob, -1)  # the built-in default; the kernel that runs ignores it
                choice[knob] = -1
                timings[knob] = "not timed: the chosen kernel does not read this knob"
                continue
            best_v, best_t = None, math.inf
            timings[knob] = {}
            for v in CANDIDATES[knob]:
                gx.tune_set(knob, v)
                t = min(run() for _ in range(repeats))
                timings[knob][str(v)] = t
                if t < best_t:
                    best_v, best_t = v, t
            gx.tune_set(knob, best_v)
            choice[knob] = best_v
    finally:
        gx.set_profiling(False)
    if cache is not None:
        table[key] = {"choice": choice, "ms": timings}
        cache.parent.mkdir(parents=True, exist_ok=True)
        cache.write_text(json.dumps(table, indent=1, sort_keys=True))
    return choice


def reset() -> None:
    """Restore every knob's built-in default."""
    import galois_amd as gx
    for knob in CANDIDATES:
        gx.tune_set(knob, -1)
