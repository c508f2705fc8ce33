"""Shared fixtures: the oracle (CPU restatement, test infrastructure) and GPU gating."""
from __future__ import annotations

import ctypes as C
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


class Oracle:
    """ctypes wrapper of oracle/liboracle.so (the checker, never the product)."""

    def __init__(self, path: Path):
        self.lib = C.CDLL(str(path))
        P, I64, U32, INT, D, F = C.c_void_p, C.c_int64, C.c_uint32, C.c_int, C.c_double, C.c_float
        sig = {
            "oracle_bfs": [I64, P, P, U32, P, INT],
            "oracle_sssp_delta": [I64, P, P, P, U32, U32, P, INT],
            "oracle_dijkstra": [I64, P, P, P, U32, P],
            "oracle_cc": [I64, P, P, P, INT],
            "oracle_kcore": [I64, P, P, U32, P, INT],
            "oracle_pagerank_fp64": [I64, P, P, D, D, U32, P, P, P, INT],
            "oracle_pagerank_residual_f32": [I64, P, P, F, F, U32, P, P, INT],
            "oracle_pagerank_topo_f32": [I64, P, P, F, F, U32, P, P, INT],
            "oracle_pagerank_push_f32": [I64, P, P, F, F, U32, P, P],
            "oracle_pagerank_weighted_fp64": [I64, P, P, P, P, D, D, U32, P, P],
            "oracle_pagerank_adjoint_fp64": [I64, P, P, P, P, P, D, D, U32, P, P],
            "oracle_pagerank_weighted_fp64_mt": [I64, P, P, P, P, D, D, U32, P, P, INT],
            "oracle_pagerank_adjoint_fp64_mt": [I64, P, P, P, P, P, D, D, U32, P, P, INT],
            "oracle_bc": [I64, P, P, P, U32, P, INT],
            "oracle_transpose": [I64, P, P, P, P, P],
        }
        for name, args in sig.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = INT

    @staticmethod
    def _p(a):
        return a.ctypes.data

    def bfs(self, row, col, src, threads=4):
        lv = np.empty(len(row) - 1, np.uint32)
        assert self.lib.oracle_bfs(len(row) - 1, self._p(row), self._p(col), src, self._p(lv), threads) == 0
        return lv

    def sssp(self, row, col, w, src, shift, threads=4):
        d = np.empty(len(row) - 1, np.uint32)
        assert self.lib.oracle_sssp_delta(len(row) - 1, self._p(row), self._p(col), self._p(w), src, shift,
                                          self._p(d), threads) == 0
        return d

    def dijkstra(self, row, col, w, src):
        d = np.empty(len(row) - 1, np.uint32)
        assert self.lib.oracle_dijkstra(len(row) - 1, self._p(row), self._p(col), self._p(w), src, self._p(d)) == 0
        return d

    def cc(self, row, col, threads=4):
        lab = np.empty(len(row) - 1, np.uint32)
        assert self.lib.oracle_cc(len(row) - 1, self._p(row), self._p(col), self._p(lab), threads) == 0
        return lab

    def kcore(self, row, col, k, threads=4):
        deg = np.empty(len(row) - 1, np.uint32)
        assert self.lib.oracle_kcore(len(row) - 1, self._p(row), self._p(col), k, self._p(deg), threads) == 0
        return deg

    def pagerank(self, row, col, alpha=0.85, tol=1e-9, max_iter=100, threads=4):
        n = len(row) - 1
        x = np.empty(n, np.float64)
        it = C.c_uint32()
        res = C.c_double()
        assert self.lib.oracle_pagerank_fp64(n, self._p(row), self._p(col), alpha, tol, max_iter, self._p(x),
                                             C.byref(it), C.byref(res), threads) == 0
        return x, it.value, res.value

    def pagerank_residual(self, row, col, alpha=0.85, tol=1e-3, max_iter=1000, threads=4):
        n = len(row) - 1
        x = np.empty(n, np.float32)
        it = C.c_uint32()
        assert self.lib.oracle_pagerank_residual_f32(n, self._p(row), self._p(col), alpha, tol, max_iter,
                                                     self._p(x), C.byref(it), threads) == 0
        return x, it.value

    def pagerank_topo(self, row, col, alpha=0.85, tol=1e-3, max_iter=1000, threads=4):
        n = len(row) - 1
        x = np.empty(n, np.float32)
        it = C.c_uint32()
        assert self.lib.oracle_pagerank_topo_f32(n, self._p(row), self._p(col), alpha, tol, max_iter,
                                                 self._p(x), C.byref(it), threads) == 0
        return x, it.value

    def pagerank_push(self, row, col, alpha=0.85, tol=1e-3, max_iter=1000):
        n = len(row) - 1
        x = np.empty(n, np.float32)
        it = C.c_uint32()
        assert self.lib.oracle_pagerank_push_f32(n, self._p(row), self._p(col), alpha, tol, max_iter,
                                                 self._p(x), C.byref(it)) == 0
        return x, it.value

    def pagerank_weighted(self, row, col, w, t, alpha=0.85, tol=1e-12, max_iter=1000, threads=1):
        n = len(row) - 1
        x = np.empty(n, np.float64)
        it = C.c_uint32()
        w = np.ascontiguousarray(w, np.float64)
        t = np.ascontiguousarray(t, np.float64)
        if threads == 1:
            assert self.lib.oracle_pagerank_weighted_fp64(n, self._p(row), self._p(col), self._p(w), self._p(t),
                                                          alpha, tol, max_iter, self._p(x), C.byref(it)) == 0
        else:
            assert self.lib.oracle_pagerank_weighted_fp64_mt(n, self._p(row), self._p(col), self._p(w), self._p(t),
                                                             alpha, tol, max_iter, self._p(x), C.byref(it),
                                                             threads) == 0
        return x

    def pagerank_adjoint(self, row, col, w, x, c, alpha=0.85, tol=1e-12, max_iter=1000, threads=1):
        n, m = len(row) - 1, len(col)
        gw = np.empty(max(m, 1), np.float64)
        gt = np.empty(n, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        c = np.ascontiguousarray(c, np.float64)
        if threads == 1:
            assert self.lib.oracle_pagerank_adjoint_fp64(n, self._p(row), self._p(col), self._p(w), self._p(x),
                                                         self._p(c), alpha, tol, max_iter, self._p(gw),
                                                         self._p(gt)) == 0
        else:
            assert self.lib.oracle_pagerank_adjoint_fp64_mt(n, self._p(row), self._p(col), self._p(w), self._p(x),
                                                            self._p(c), alpha, tol, max_iter, self._p(gw),
                                                            self._p(gt), threads) == 0
        return gw[:m], gt

    def bc(self, row, col, sources, threads=4):
        n = len(row) - 1
        src = np.ascontiguousarray(sources, np.uint32)
        out = np.empty(n, np.float64)
        assert self.lib.oracle_bc(n, self._p(row), self._p(col), self._p(src), len(src), self._p(out),
                                  threads) == 0
        return out

    def transpose(self, row, col):
        n, m = len(row) - 1, len(col)
        trow = np.empty(n + 1, np.int32)
        tcol = np.empty(max(m, 1), np.int32)
        perm = np.empty(max(m, 1), np.int32)
        assert self.lib.oracle_transpose(n, self._p(row), self._p(col), self._p(trow), self._p(tcol),
                                         self._p(perm)) == 0
        return trow, tcol[:m], perm[:m]


@pytest.fixture(scope="session")
def oracle() -> Oracle:
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
    return Oracle(ORACLE_SO)


@pytest.fixture(scope="session")
def golden() -> dict:
    with open(GOLDEN / "fixtures.json") as f:
        return json.load(f)


@pytest.fixture(scope="session")
def rmat16(gx):
    """Symmetric weighted RMAT-16 generated on the device, with its host CSR."""
    print("generate rmat16 symmetric weighted", flush=True)
    G = gx.Graph.rmat(16, 16, symmetric=True, weights="uniform", wlo=1, whi=255)
    row, col, w = G.to_host()
    return G, row, col, w


def csr_arrays(g: dict):
    row = np.asarray(g["row"], dtype=np.int32)
    col = np.asarray(g["col"], dtype=np.int32)
    w = np.asarray(g["w"], dtype=np.uint32) if g.get("w") is not None else None
    return row, col, w


@pytest.fixture(scope="session")
def gx():
    """The product package; GPU tests only."""
    import galois_amd
    from galois_amd import _lib
    _lib.check(_lib.lib().gx_init(0))
    return galois_amd
