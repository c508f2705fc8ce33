"""Writes tests/golden/fixtures.json: tiny graphs with expected outputs.

Expected values come from libraries and methods independent of both the
oracle and the HIP library:
  - BFS levels: networkx.single_source_shortest_path_length
  - SSSP distances: scipy.sparse.csgraph.dijkstra
  - CC labels: scipy.sparse.csgraph.connected_components relabelled to the
    minimum vertex id of each component
  - k-core membership: networkx.k_core
  - BC: networkx.betweenness_centrality(DiGraph, normalized=False) over all
    sources, and brute-force enumeration of all shortest paths
    (networkx.all_shortest_paths) for a source subset
  - PageRank: dense direct solves (numpy.linalg.solve) of the normalised
    (dangling-aware) and the unnormalised fixed points
  - differentiable PageRank: dense solve of the forward system and central
    finite differences of the dense solve for the gradients
Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
from pathlib import Path

import networkx as nx
import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

OUT = Path(__file__).resolve().parent / "fixtures.json"
INF = 2147483646
SEED = 27491095


def csr_from_edges(n, edges, symmetric):
    es = set()
    for u, v in edges:
        if u == v:
            continue
        es.add((u, v))
        if symmetric:
            es.add((v, u))
    es = sorted(es)
    row = np.zeros(n + 1, dtype=np.int64)
    for u, _ in es:
        row[u + 1] += 1
    row = np.cumsum(row)
    col = [v for _, v in es]
    return row.tolist(), col


def weights_for(row, col, rng, symmetric, lo=1, hi=20):
    n = len(row) - 1
    wmap = {}
    w = []
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            v = col[e]
            key = (min(u, v), max(u, v)) if symmetric else (u, v)
            if key not in wmap:
                wmap[key] = int(rng.integers(lo, hi + 1))
            w.append(wmap[key])
    return w


def to_nx(n, row, col, directed):
    G = nx.DiGraph() if directed else nx.Graph()
    G.add_nodes_from(range(n))
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            G.add_edge(u, col[e])
    return G


def scipy_csr(n, row, col, w=None):
    data = np.ones(len(col)) if w is None else np.asarray(w, dtype=np.float64)
    return sp.csr_matrix((data, np.asarray(col), np.asarray(row)), shape=(n, n))


def bfs_levels(G, n, s):
    lv = [INF] * n
    for v, d in nx.single_source_shortest_path_length(G, s).items():
        lv[v] = d
    return lv


def sssp(n, row, col, w, s):
    d = csg.dijkstra(scipy_csr(n, row, col, w), directed=True, indices=s)
    return [INF if not np.isfinite(x) else int(round(x)) for x in d]


def cc_min_labels(n, row, col):
    _, lab = csg.connected_components(scipy_csr(n, row, col), directed=False)
    mins = {}
    for v in range(n):
        mins.setdefault(lab[v], v)
    return [int(mins[lab[v]]) for v in range(n)]


def kcore(G, n, row, col, k):
    core = set(nx.k_core(G, k).nodes()) if G.number_of_nodes() else set()
    alive = [1 if v in core else 0 for v in range(n)]
    deg = [sum(1 for e in range(row[v], row[v + 1]) if col[e] in core) for v in range(n)]
    return alive, deg


def bc_subset_bruteforce(G, n, sources):
    bc = [0.0] * n
    for s in sources:
        for t in range(n):
            if t == s or not nx.has_path(G, s, t):
                continue
            paths = list(nx.all_shortest_paths(G, s, t))
            for p in paths:
                for v in p[1:-1]:
                    bc[v] += 1.0 / len(paths)
    return bc


def pagerank_normalised(n, row, col, alpha=0.85):
    outdeg = np.diff(np.asarray(row))
    M = np.zeros((n, n))
    for u in range(n):
        if outdeg[u] == 0:
            M[:, u] += 1.0 / n
        for e in range(row[u], row[u + 1]):
            M[col[e], u] += 1.0 / outdeg[u]
    x = np.linalg.solve(np.eye(n) - alpha * M, np.full(n, (1 - alpha) / n))
    return x.tolist()


def pagerank_unnormalised(n, row, col, alpha=0.85):
    outdeg = np.diff(np.asarray(row))
    M = np.zeros((n, n))
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            M[col[e], u] += 1.0 / outdeg[u]
    x = np.linalg.solve(np.eye(n) - alpha * M, np.full(n, 1 - alpha))
    return x.tolist()


def weighted_forward(n, row, col, w, t, alpha):
    outdeg = np.diff(np.asarray(row))
    A = np.zeros((n, n))
    for u in range(n):
        if outdeg[u] == 0:
            A[:, u] += 1.0 / n
            continue
        s = sum(w[row[u]:row[u + 1]])
        for e in range(row[u], row[u + 1]):
            A[col[e], u] += w[e] / s
    return np.linalg.solve(np.eye(n) - alpha * A, (1 - alpha) * np.asarray(t))


def weighted_case(rng):
    n = 12
    edges = [(int(rng.integers(n)), int(rng.integers(n))) for _ in range(30)]
    row, col = csr_from_edges(n, edges, symmetric=False)
    w = np.exp(rng.normal(size=len(col))).tolist()
    t = rng.random(n)
    t = (t / t.sum()).tolist()
    c = rng.normal(size=n).tolist()
    alpha = 0.85
    x = weighted_forward(n, row, col, w, t, alpha)
    eps = 1e-6
    gw = []
    for e in range(len(col)):
        wp = list(w)
        wm = list(w)
        wp[e] += eps
        wm[e] -= eps
        gw.append(float((np.dot(c, weighted_forward(n, row, col, wp, t, alpha))
                         - np.dot(c, weighted_forward(n, row, col, wm, t, alpha))) / (2 * eps)))
    gt = []
    for v in range(n):
        tp = list(t)
        tm = list(t)
        tp[v] += eps
        tm[v] -= eps
        gt.append(float((np.dot(c, weighted_forward(n, row, col, w, tp, alpha))
                         - np.dot(c, weighted_forward(n, row, col, w, tm, alpha))) / (2 * eps)))
    return {"name": "weighted12", "n": n, "row": row, "col": col, "w": w, "t": t, "c": c, "alpha": alpha,
            "x": x.tolist(), "grad_w": gw, "grad_t": gt}


def undirected_cases(rng):
    cases = []
    cases.append(("path10", 10, [(i, i + 1) for i in range(9)]))
    cases.append(("star10", 10, [(0, i) for i in range(1, 10)]))
    grid = [(r * 5 + c, r * 5 + c + 1) for r in range(5) for c in range(4)]
    grid += [(r * 5 + c, (r + 1) * 5 + c) for r in range(4) for c in range(5)]
    cases.append(("grid5x5", 25, grid))
    cases.append(("components11", 11, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (5, 6), (6, 7), (7, 8)]))
    G = nx.gnp_random_graph(40, 0.12, seed=SEED)
    cases.append(("er40", 40, list(G.edges())))
    G = nx.barabasi_albert_graph(60, 2, seed=SEED)
    cases.append(("ba60", 60, list(G.edges())))
    G = nx.complete_graph(6)
    cases.append(("k6", 6, list(G.edges())))
    return cases


def main():
    rng = np.random.default_rng(SEED)
    out = {"undirected": [], "directed": [], "weighted": []}
    for name, n, edges in undirected_cases(rng):
        row, col = csr_from_edges(n, edges, symmetric=True)
        w = weights_for(row, col, rng, symmetric=True)
        G = to_nx(n, row, col, directed=False)
        D = to_nx(n, row, col, directed=True)
        deg = np.diff(row)
        srcs = sorted({0, int(np.argmax(deg)), n - 1})
        case = {"name": name, "n": n, "row": row, "col": col, "w": w,
                "bfs": {str(s): bfs_levels(G, n, s) for s in srcs},
                "sssp": {str(s): sssp(n, row, col, w, s) for s in srcs},
                "cc": cc_min_labels(n, row, col),
                "kcore": {str(k): dict(zip(("alive", "degree"), kcore(G, n, row, col, k))) for k in (2, 3)},
                "bc_all": [float(x) for x in
                           (lambda d: [d[v] for v in range(n)])(nx.betweenness_centrality(D, normalized=False))],
                "bc_sources": srcs,
                "bc_subset": bc_subset_bruteforce(D, n, srcs),
                "pagerank": pagerank_normalised(n, row, col),
                "pagerank_unnormalised": pagerank_unnormalised(n, row, col)}
        out["undirected"].append(case)
    for name, n, p in (("dir30", 30, 0.08), ("dir50", 50, 0.05)):
        G = nx.gnp_random_graph(n, p, seed=SEED + n, directed=True)
        row, col = csr_from_edges(n, list(G.edges()), symmetric=False)
        D = to_nx(n, row, col, directed=True)
        out["directed"].append({"name": name, "n": n, "row": row, "col": col,
                                "bfs0": bfs_levels(D, n, 0),
                                "pagerank": pagerank_normalised(n, row, col),
                                "pagerank_unnormalised": pagerank_unnormalised(n, row, col)})
    out["weighted"].append(weighted_case(rng))
    OUT.write_text(json.dumps(out, separators=(",", ":")))
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
