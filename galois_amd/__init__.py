"""galois_amd — lonestar analytics (BFS, SSSP, CC, k-core, PageRank, BC) on MI355X.

Python host mirror of the C ABI in include/galois_amd.h.  Graphs live in
device memory (`Graph`); every algorithm writes a caller-visible device
tensor (torch, on the current GPU) and returns it together with the call's
statistics.  All computation happens in libgalois_amd.so's HIP kernels;
PyTorch only provides device memory.

The function names follow the reference's Python bindings
(python/galois/_bfs.pyx:171 `bfs`, _sssp.pyx:154 `sssp`, _pagerank.pyx:157
`pagerank`, _connected_components.pyx:87 `connectedComponents`), taking a
Graph instead of (numThreads, filename).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (GX_BC_INFINITY, GX_DIST_INFINITY, GX_GRAPH_FWEIGHTED, GX_GRAPH_SYMMETRIC,
                   GX_GRAPH_WEIGHTED, GxError, check, gx_host_csr, gx_stats, lib)

__all__ = [
    "Graph", "GxError", "Stats", "bfs", "bfs_do", "sssp", "connected_components", "connectedComponents",
    "kcore", "pagerank", "pagerank_ref", "pagerank_push", "pagerank_weighted", "pagerank_adjoint",
    "betweenness_centrality", "betweenness_centrality_batched", "pick_sources", "degree_power_law", "read_gr",
    "write_gr",
    "GX_DIST_INFINITY", "GX_BC_INFINITY", "set_profiling", "differentiable_pagerank",
]

KRAND_SEED = 27491095  # lonestar/liblonestar/include/Lonestar/Utils.h:31


def _torch():
    import torch
    return torch


def _ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


@dataclass
class Stats:
    iterations: int
    launches: int
    edges_visited: int
    items: int
    residual: float
    elapsed_ms: float
    main_kernel_ms: float
    main_kernel_launches: int
    main_kernel_bytes: int


def _stats(st: gx_stats) -> Stats:
    return Stats(**st.as_dict())


def set_profiling(on: bool) -> None:
    """Time every launch of each entry point's dominant kernel with HIP events."""
    lib().gx_set_profiling(1 if on else 0)


def tune_set(key: str, value: int) -> None:
    """Set a tuning knob (gx_tune_set); -1 restores the built-in default."""
    check(lib().gx_tune_set(key.encode(), int(value)))


def tune_get(key: str) -> int:
    """Current value of a tuning knob; -1 means the built-in default."""
    v = C.c_int64()
    check(lib().gx_tune_get(key.encode(), C.byref(v)))
    return v.value


def tune_effective(key: str, n: int, m: int) -> int:
    """Value of a knob that an entry point uses on a graph of n vertices and m
    edges: the value set, else the built-in default for that size."""
    v = C.c_int64()
    check(lib().gx_tune_effective(key.encode(), int(n), int(m), C.byref(v)))
    return v.value


class Graph:
    """A CSR graph resident in device memory (owns a gx_graph handle)."""

    def __init__(self, handle: int, keepalive=None):
        self._h = C.c_void_p(handle)
        self._keepalive = keepalive
        n, m, fl = C.c_int64(), C.c_int64(), C.c_uint32()
        check(lib().gx_graph_info(self._h, C.byref(n), C.byref(m), C.byref(fl)))
        self.n, self.m, self.flags = n.value, m.value, fl.value

    # ------------------------------------------------------------ builders
    @classmethod
    def from_csr(cls, row, col, weights=None, symmetric: bool = False) -> "Graph":
        """Copy host CSR arrays (numpy) to the device."""
        row = np.ascontiguousarray(row, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        n, m = len(row) - 1, len(col)
        flags = GX_GRAPH_SYMMETRIC if symmetric else 0
        wp = None
        if weights is not None:
            weights = np.ascontiguousarray(weights)
            if weights.dtype == np.float32:
                flags |= GX_GRAPH_FWEIGHTED
            else:
                weights = weights.astype(np.uint32)
                flags |= GX_GRAPH_WEIGHTED
            wp = weights.ctypes.data
        h = C.c_void_p()
        check(lib().gx_graph_create(n, m, row.ctypes.data, col.ctypes.data if m else None, wp, flags,
                                    C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = KRAND_SEED, symmetric: bool = True, weights: str | None = None, wlo: int = 1,
             whi: int = 255) -> "Graph":
        """Seeded RMAT graph generated on the device (self-loops/duplicates removed)."""
        kind = {None: 0, "uniform": 1, "lognormal": 2}[weights]
        h = C.c_void_p()
        check(lib().gx_gen_rmat(scale, edge_factor, a, b, c, seed, int(symmetric), kind, wlo, whi, C.byref(h)))
        return cls(h.value)

    @classmethod
    def grid(cls, side: int, drop_ppm: int = 50000, shortcut_ppm: int = 10000, seed: int = KRAND_SEED,
             wlo: int = 1, whi: int = 1000) -> "Graph":
        """Road-like side x side grid (symmetric, uint32 weights) generated on the device."""
        h = C.c_void_p()
        check(lib().gx_gen_grid(side, drop_ppm, shortcut_ppm, seed, wlo, whi, C.byref(h)))
        return cls(h.value)

    @classmethod
    def union(cls, parts: list["Graph"]) -> tuple["Graph", np.ndarray]:
        """Disjoint union (one CSR) of several graphs and the vertex offsets."""
        arr = (C.c_void_p * len(parts))(*[p._h.value for p in parts])
        offs = np.zeros(len(parts) + 1, dtype=np.int64)
        h = C.c_void_p()
        check(lib().gx_graph_union(arr, len(parts), C.byref(h), offs.ctypes.data))
        return cls(h.value), offs

    @classmethod
    def from_device(cls, row, col, weights=None, symmetric: bool = False) -> "Graph":
        """Wrap device tensors (torch int32) without copying; they are kept alive."""
        flags = GX_GRAPH_SYMMETRIC if symmetric else 0
        torch = _torch()
        if weights is not None:
            flags |= GX_GRAPH_FWEIGHTED if weights.dtype == torch.float32 else GX_GRAPH_WEIGHTED
        h = C.c_void_p()
        check(lib().gx_graph_create_device(row.numel() - 1, col.numel(), _ptr(row), _ptr(col), _ptr(weights),
                                           flags, C.byref(h)))
        return cls(h.value, keepalive=(row, col, weights))

    # ------------------------------------------------------------- queries
    @property
    def symmetric(self) -> bool:
        return bool(self.flags & GX_GRAPH_SYMMETRIC)

    @property
    def weighted(self) -> bool:
        return bool(self.flags & (GX_GRAPH_WEIGHTED | GX_GRAPH_FWEIGHTED))

    def to_host(self):
        """(row int32[n+1], col int32[m], weights uint32/float32[m] or None) as numpy arrays."""
        row = np.empty(self.n + 1, dtype=np.int32)
        col = np.empty(max(self.m, 1), dtype=np.int32)
        w = None
        if self.flags & GX_GRAPH_WEIGHTED:
            w = np.empty(max(self.m, 1), dtype=np.uint32)
        elif self.flags & GX_GRAPH_FWEIGHTED:
            w = np.empty(max(self.m, 1), dtype=np.float32)
        check(lib().gx_graph_download(self._h, row.ctypes.data, col.ctypes.data,
                                      None if w is None else w.ctypes.data))
        return row, col[:self.m], (None if w is None else w[:self.m])

    def row_host(self) -> np.ndarray:
        """Row offsets int32[n+1] only (the edge arrays are not copied)."""
        row = np.empty(self.n + 1, dtype=np.int32)
        check(lib().gx_graph_download(self._h, row.ctypes.data, None, None))
        return row

    def transpose_host(self):
        """(trow, tcol, perm) of the device-built transpose as numpy arrays."""
        trow = np.empty(self.n + 1, dtype=np.int32)
        tcol = np.empty(max(self.m, 1), dtype=np.int32)
        perm = np.empty(max(self.m, 1), dtype=np.int32)
        check(lib().gx_graph_download_transpose(self._h, trow.ctypes.data, tcol.ctypes.data, perm.ctypes.data))
        return trow, tcol[:self.m], perm[:self.m]

    def build_transpose(self) -> None:
        check(lib().gx_graph_build_transpose(self._h))

    def set_stream(self, stream) -> None:
        """Run this graph's work on a torch.cuda.Stream (or raw hipStream_t int)."""
        raw = getattr(stream, "cuda_stream", stream)
        check(lib().gx_graph_set_stream(self._h, raw))

    def free(self) -> None:
        if self._h is not None and self._h.value:
            lib().gx_graph_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------- .gr files
def read_gr(path: str):
    """Read a .gr file into host arrays: (row, col, edge_data bytes view or None, edge_size)."""
    h = gx_host_csr()
    check(lib().gx_gr_read(str(path).encode(), C.byref(h)))
    try:
        n, m = h.n, h.m
        row = np.ctypeslib.as_array(h.row, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(h.col, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
        data = None
        if h.edge_size and m:
            raw = (C.c_uint8 * (m * h.edge_size)).from_address(h.edge_data)
            data = np.frombuffer(bytes(raw), dtype=np.uint8)
        return row, col, data, h.edge_size
    finally:
        lib().gx_host_csr_free(C.byref(h))


def write_gr(path: str, row, col, edge_data=None, edge_size: int = 0) -> None:
    """Write host CSR arrays as a version-1 .gr file."""
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    h = gx_host_csr()
    h.n, h.m = len(row) - 1, len(col)
    h.row = row.ctypes.data_as(C.POINTER(C.c_int32))
    h.col = col.ctypes.data_as(C.POINTER(C.c_int32))
    if edge_data is not None:
        edge_data = np.ascontiguousarray(edge_data)
        h.edge_data = edge_data.ctypes.data
        h.edge_size = edge_size or edge_data.itemsize
    check(lib().gx_gr_write(str(path).encode(), C.byref(h)))


def pick_sources(row, count: int, seed: int = KRAND_SEED) -> np.ndarray:
    """lonestar SourcePicker (Utils.h:12-31) over a host row-offset array."""
    row = np.ascontiguousarray(row, dtype=np.int32)
    out = np.empty(max(count, 1), dtype=np.uint32)
    check(lib().gx_pick_sources(len(row) - 1, row.ctypes.data, count, seed, out.ctypes.data))
    return out[:count]


def degree_power_law(row) -> bool:
    """The reference's AutoAlgo degree test (Utils.h:37-57) on a host CSR row
    array: average degree >= 10 and, over 1000 picked vertices, mean / 1.25 >
    median (gx_degree_power_law)."""
    row = np.ascontiguousarray(row, dtype=np.int32)
    out = C.c_int()
    check(lib().gx_degree_power_law(len(row) - 1, int(row[-1]), row.ctypes.data, C.byref(out)))
    return bool(out.value)


# -------------------------------------------------------------- algorithms
def _empty(n, dtype):
    torch = _torch()
    return torch.empty(max(n, 1), dtype=dtype, device="cuda")


def bfs(g: Graph, source: int):
    """Level BFS; returns (level int32 tensor viewed as uint32 values, Stats)."""
    torch = _torch()
    level = _empty(g.n, torch.int32)
    st = gx_stats()
    check(lib().gx_bfs(g.handle, source, _ptr(level), C.byref(st)))
    return level[:g.n], _stats(st)


def bfs_do(g: Graph, source: int, alpha: int = 15, beta: int = 18, parents: bool = False):
    """Direction-optimising BFS; returns (level, parent or None, Stats)."""
    torch = _torch()
    level = _empty(g.n, torch.int32)
    parent = _empty(g.n, torch.int32) if parents else None
    st = gx_stats()
    check(lib().gx_bfs_do(g.handle, source, alpha, beta, _ptr(level), _ptr(parent), C.byref(st)))
    return level[:g.n], (None if parent is None else parent[:g.n]), _stats(st)


def sssp(g: Graph, source: int, shift: int = 13):
    """Delta-stepping SSSP (bucket = dist >> shift); returns (dist, Stats)."""
    torch = _torch()
    dist = _empty(g.n, torch.int32)
    st = gx_stats()
    check(lib().gx_sssp_delta(g.handle, source, shift, _ptr(dist), C.byref(st)))
    return dist[:g.n], _stats(st)


def sssp_batched(g: Graph, sources, shift: int = 13):
    """Shortest paths from every source at once (64 per wavefront); returns
    (dist int32 tensor [len(sources), n], Stats), row i for sources[i]."""
    torch = _torch()
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    dist = torch.empty((len(src), g.n), dtype=torch.int32, device="cuda")
    st = gx_stats()
    check(lib().gx_sssp_batched(g.handle, src.ctypes.data, len(src), shift, _ptr(dist), C.byref(st)))
    return dist, _stats(st)


def connected_components(g: Graph):
    """Min-label connected components; returns (label, Stats)."""
    torch = _torch()
    label = _empty(g.n, torch.int32)
    st = gx_stats()
    check(lib().gx_cc(g.handle, _ptr(label), C.byref(st)))
    return label[:g.n], _stats(st)


connectedComponents = connected_components


def kcore(g: Graph, k: int):
    """k-core peeling; returns (degree, alive uint8, alive_count, Stats)."""
    torch = _torch()
    deg = _empty(g.n, torch.int32)
    alive = _empty(g.n, torch.uint8)
    cnt = C.c_uint64()
    st = gx_stats()
    check(lib().gx_kcore(g.handle, k, _ptr(deg), _ptr(alive), C.byref(cnt), C.byref(st)))
    return deg[:g.n], alive[:g.n], cnt.value, _stats(st)


def pagerank(g: Graph, alpha: float = 0.85, tol: float = 1e-9, max_iter: int = 100, precision: str = "fp32"):
    """Normalised pull PageRank; precision 'fp32' or 'fp64'.  FP32 row sums are exact 64-bit fixed-point
    sums in the default propagation-blocked sweep; the merge-path pull (pr.blocking = 0) uses Kahan sums.
    Returns (pr, Stats)."""
    torch = _torch()
    fp64 = precision == "fp64"
    pr = _empty(g.n, torch.float64 if fp64 else torch.float32)
    st = gx_stats()
    check(lib().gx_pagerank_pull(g.handle, alpha, tol, max_iter, 1 if fp64 else 0, _ptr(pr), C.byref(st)))
    return pr[:g.n], _stats(st)


def pagerank_layout_info(g: Graph) -> dict:
    """Shape of the propagation-blocked layout of g (empty dict before the
    first FP32 pagerank call builds it)."""
    info = _lib.gx_pb_layout_info()
    check(lib().gx_pagerank_layout_info(g.handle, C.byref(info)))
    return info.as_dict() if info.built else {}


def pagerank_ref(g: Graph, variant: str = "residual", alpha: float = 0.85, tol: float = 1e-3,
                 max_iter: int = 1000):
    """Reference-semantics (unnormalised) FP32 pull PageRank: 'residual' or 'topo'."""
    torch = _torch()
    pr = _empty(g.n, torch.float32)
    st = gx_stats()
    check(lib().gx_pagerank_pull_ref(g.handle, {"residual": 0, "topo": 1}[variant], alpha, tol, max_iter,
                                     _ptr(pr), C.byref(st)))
    return pr[:g.n], _stats(st)


def pagerank_push(g: Graph, alpha: float = 0.85, tol: float = 1e-3, max_iter: int = 1000):
    """Synchronous residual-push PageRank (reference semantics, FP32)."""
    torch = _torch()
    pr = _empty(g.n, torch.float32)
    st = gx_stats()
    check(lib().gx_pagerank_push(g.handle, alpha, tol, max_iter, _ptr(pr), C.byref(st)))
    return pr[:g.n], _stats(st)


def pagerank_weighted(g: Graph, weights, teleport, alpha: float = 0.85, tol: float = 1e-9, max_iter: int = 200,
                      precision: str = "fp32"):
    """Forward of the differentiable PageRank: weights (float32 device tensor, CSR
    edge order, or None for the graph's own) and teleport (float32 device
    tensor).  Returns (x, Stats); x is float32 or float64 per `precision`."""
    torch = _torch()
    fp64 = precision == "fp64"
    x = _empty(g.n, torch.float64 if fp64 else torch.float32)
    st = gx_stats()
    check(lib().gx_pagerank_weighted(g.handle, _ptr(weights), _ptr(teleport), alpha, tol, max_iter, int(fp64),
                                     _ptr(x), C.byref(st)))
    return x[:g.n], _stats(st)


def pagerank_adjoint(g: Graph, weights, x, c, alpha: float = 0.85, tol: float = 1e-9, max_iter: int = 200,
                     precision: str = "fp32"):
    """Adjoint of pagerank_weighted for L with dL/dx = c: returns (grad_w, grad_t, Stats)."""
    torch = _torch()
    fp64 = precision == "fp64"
    dt = torch.float64 if fp64 else torch.float32
    x = x.to(dt).contiguous()
    gw = _empty(g.m, dt)
    gt = _empty(g.n, dt)
    st = gx_stats()
    check(lib().gx_pagerank_adjoint(g.handle, _ptr(weights), _ptr(x), _ptr(c.float().contiguous()), alpha, tol,
                                    max_iter, int(fp64), _ptr(gw), _ptr(gt), C.byref(st)))
    return gw[:g.m], gt[:g.n], _stats(st)


def _bc(fn, g: Graph, sources):
    torch = _torch()
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    bc = _empty(g.n, torch.float64)
    st = gx_stats()
    check(fn(g.handle, src.ctypes.data, len(src), _ptr(bc), C.byref(st)))
    return bc[:g.n], _stats(st)


def betweenness_centrality(g: Graph, sources):
    """Brandes BC over the given sources, one source per pass; returns (bc float64, Stats)."""
    return _bc(lib().gx_bc, g, sources)


def betweenness_centrality_batched(g: Graph, sources):
    """Brandes BC with all sources batched (64 per wavefront); returns (bc float64, Stats)."""
    return _bc(lib().gx_bc_batched, g, sources)


def differentiable_pagerank(g: Graph, weights, teleport, alpha: float = 0.85, tol: float = 1e-9,
                            max_iter: int = 200, precision: str = "fp32"):
    """PageRank as a torch.autograd function of the edge weights and the teleport
    vector (see galois_amd.autograd)."""
    from .autograd import pagerank_fn
    return pagerank_fn(g, weights, teleport, alpha, tol, max_iter, precision)
