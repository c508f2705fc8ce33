"""D-IrGL personality partition (the gx_part_* entry points of the C ABI)
and the host loops of the reference's D-IrGL apps for one host.

The loops follow lonestar/analytics/distributed/bfs/bfs_push.cpp:95-300
(InitializeGraph over all nodes, FirstItr_BFS on the source, BFS rounds over
the nodes with edges until no vertex is active), the same structure in
sssp/sssp_push.cpp:95-300 (weights pushed) and connected-components/
cc_push.cpp:85-230 (labels pushed, FirstItr over the nodes with edges), and
pagerank/pagerank_push.cpp:110-300 (ResetGraph, InitializeGraph, then
PageRank_delta + PageRank rounds until no node pushes),
betweennesscentrality/bc_level.cpp:440-720 (per source: InitializeIteration,
forward passes until no update, then MiddleSync, backward passes from the
deepest level and the accumulation into the centrality), and
lonestar/analytics/distributed/k-core/kcore_push.cpp:93-318
(InitializeGraph1 over all nodes, InitializeGraph2 over the nodes with edges,
KCoreStep1 + KCoreStep2 rounds until no vertex dies).  With one host there is
no gluon synchronisation between the operators; every other step is the
reference's."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

DIST_CURRENT, DIST_OLD, CURRENT_DEGREE, TRIM, FLAG, COMP_CURRENT, COMP_OLD = range(7)
PR_VALUE, PR_NOUT, PR_RESIDUAL, PR_DELTA = range(7, 11)
BC_LENGTH, BC_DEPENDENCY, BC_VALUE, BC_PATHS = range(11, 15)
PR_ALPHA = 1.0 - 0.85  # pagerank_push.cpp:68 (reset probability)


class PrSanity(C.Structure):
    """gx_pr_sanity: PageRankSanity's reductions."""
    _fields_ = [("max_value", C.c_float), ("min_value", C.c_float), ("max_residual", C.c_float),
                ("min_residual", C.c_float), ("sum_value", C.c_double), ("sum_residual", C.c_double),
                ("residual_over_tolerance", C.c_uint64)]
RANGE_ALL, RANGE_MASTERS, RANGE_WITH_EDGES = range(3)
INFINITY = 0xFFFFFFFF // 4  # bfs_push.cpp:72 (numeric_limits<uint32_t>::max() / 4)


class Partition:
    """One host's partition: the CSR of its local nodes (uint32 row_start and
    edge_dst, as MarshalGraph), local-to-global ids, the master range and the
    count of nodes with edges (default: up to the last node with out-edges)."""

    def __init__(self, row, col, gid=None, num_owned=None, begin_master=0, num_with_edges=None, weights=None,
                 device=0):
        h = C.c_void_p()
        check(lib().gx_part_create(device, C.byref(h)))
        self._h = h.value
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        self.n, self.m = len(row) - 1, len(col)
        if num_owned is None:
            num_owned = self.n - begin_master
        if num_with_edges is None:
            nz = np.nonzero(np.diff(row.astype(np.int64)))[0]
            num_with_edges = int(nz[-1]) + 1 if len(nz) else 0
        self.gid = None if gid is None else np.ascontiguousarray(gid, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        check(lib().gx_part_load(self._h, self.n, self.m, row.ctypes.data, col.ctypes.data if self.m else None,
                                 None if w is None or not self.m else w.ctypes.data,
                                 None if self.gid is None else self.gid.ctypes.data, num_owned, begin_master,
                                 num_with_edges))

    def free(self) -> None:
        if getattr(self, "_h", None):
            lib().gx_part_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass

    # ------------------------------------------------------------- ranges
    def range(self, which: int) -> tuple[int, int]:
        b, e = C.c_uint32(), C.c_uint32()
        check(lib().gx_part_range(self._h, which, C.byref(b), C.byref(e)))
        return b.value, e.value

    # ------------------------------------------------------------- fields
    def get(self, field: int, lid: int) -> int:
        v = C.c_uint32()
        check(lib().gx_part_get(self._h, field, lid, C.byref(v)))
        return v.value

    def set(self, field: int, lid: int, value: int) -> None:
        check(lib().gx_part_set(self._h, field, lid, value))

    def add(self, field: int, lid: int, value: int) -> None:
        check(lib().gx_part_add(self._h, field, lid, value))

    def min(self, field: int, lid: int, value: int) -> bool:
        changed = C.c_int()
        check(lib().gx_part_min(self._h, field, lid, value, C.byref(changed)))
        return bool(changed.value)

    def fill(self, field: int, begin: int, end: int, value: int) -> None:
        check(lib().gx_part_fill(self._h, field, begin, end, value))

    def read(self, field: int, begin: int = 0, end: int | None = None) -> np.ndarray:
        end = self.n if end is None else end
        out = np.empty(max(end - begin, 1), dtype=np.uint32)
        check(lib().gx_part_read(self._h, field, begin, end, out.ctypes.data))
        return out[:end - begin]

    def bitset(self, field: int) -> np.ndarray:
        words = np.empty((self.n + 63) // 64, dtype=np.uint64)
        check(lib().gx_part_bitset_get(self._h, field, words.ctypes.data))
        return words

    def bitset_reset(self, field: int, begin: int = 0, end: int | None = None) -> None:
        check(lib().gx_part_bitset_reset(self._h, field, begin, self.n if end is None else end))

    # ---------------------------------------------------------- operators
    def bfs_init(self, begin: int, end: int, infinity: int, src_gid: int) -> None:
        check(lib().gx_part_bfs_init(self._h, begin, end, infinity, src_gid))

    def bfs_first(self, begin: int, end: int) -> None:
        check(lib().gx_part_bfs_first(self._h, begin, end))

    def bfs(self, begin: int, end: int, priority: int) -> tuple[int, int]:
        a, w = C.c_uint32(), C.c_uint32()
        check(lib().gx_part_bfs(self._h, begin, end, priority, C.byref(a), C.byref(w)))
        return a.value, w.value

    def bfs_check(self, begin: int, end: int, infinity: int) -> tuple[int, int]:
        s, mx = C.c_uint64(), C.c_uint32()
        check(lib().gx_part_bfs_check(self._h, begin, end, infinity, C.byref(s), C.byref(mx)))
        return s.value, mx.value

    def sssp_init(self, begin: int, end: int, infinity: int, src_gid: int) -> None:
        check(lib().gx_part_sssp_init(self._h, begin, end, infinity, src_gid))

    def sssp_first(self, begin: int, end: int) -> None:
        check(lib().gx_part_sssp_first(self._h, begin, end))

    def sssp(self, begin: int, end: int, priority: int) -> tuple[int, int]:
        a, w = C.c_uint32(), C.c_uint32()
        check(lib().gx_part_sssp(self._h, begin, end, priority, C.byref(a), C.byref(w)))
        return a.value, w.value

    def sssp_check(self, begin: int, end: int, infinity: int) -> tuple[int, int, int]:
        s, d, mx = C.c_uint64(), C.c_uint64(), C.c_uint32()
        check(lib().gx_part_sssp_check(self._h, begin, end, infinity, C.byref(s), C.byref(d), C.byref(mx)))
        return s.value, d.value, mx.value

    def cc_init(self, begin: int, end: int) -> None:
        check(lib().gx_part_cc_init(self._h, begin, end))

    def cc_first(self, begin: int, end: int) -> None:
        check(lib().gx_part_cc_first(self._h, begin, end))

    def cc(self, begin: int, end: int) -> int:
        a = C.c_uint32()
        check(lib().gx_part_cc(self._h, begin, end, C.byref(a)))
        return a.value

    def cc_check(self, begin: int, end: int) -> int:
        a = C.c_uint64()
        check(lib().gx_part_cc_check(self._h, begin, end, C.byref(a)))
        return a.value

    def pr_reset(self, begin: int, end: int) -> None:
        check(lib().gx_part_pr_reset(self._h, begin, end))

    def pr_init(self, begin: int, end: int, alpha: float = PR_ALPHA) -> None:
        check(lib().gx_part_pr_init(self._h, begin, end, alpha))

    def pr_delta(self, begin: int, end: int, alpha: float, tolerance: float) -> None:
        check(lib().gx_part_pr_delta(self._h, begin, end, alpha, tolerance))

    def pr(self, begin: int, end: int) -> int:
        a = C.c_uint32()
        check(lib().gx_part_pr(self._h, begin, end, C.byref(a)))
        return a.value

    def pr_check(self, begin: int, end: int, tolerance: float) -> PrSanity:
        out = PrSanity()
        check(lib().gx_part_pr_check(self._h, begin, end, tolerance, C.byref(out)))
        return out

    def read_float(self, field: int, begin: int = 0, end: int | None = None) -> np.ndarray:
        return self.read(field, begin, end).view(np.float32)

    def bc_init(self, begin: int, end: int) -> None:
        check(lib().gx_part_bc_init(self._h, begin, end))

    def bc_iter(self, begin: int, end: int, infinity: int, src_gid: int) -> None:
        check(lib().gx_part_bc_iter(self._h, begin, end, infinity, src_gid))

    def bc_forward(self, begin: int, end: int, round_: int) -> int:
        d = C.c_uint32()
        check(lib().gx_part_bc_forward(self._h, begin, end, round_, C.byref(d)))
        return d.value

    def bc_middle(self, begin: int, end: int, infinity: int) -> None:
        check(lib().gx_part_bc_middle(self._h, begin, end, infinity))

    def bc_backward(self, begin: int, end: int, round_: int) -> None:
        check(lib().gx_part_bc_backward(self._h, begin, end, round_))

    def bc_accumulate(self, begin: int, end: int) -> None:
        check(lib().gx_part_bc_accumulate(self._h, begin, end))

    def bc_check(self, begin: int, end: int) -> tuple[float, float, float]:
        s, mx, mn = C.c_float(), C.c_float(), C.c_float()
        check(lib().gx_part_bc_check(self._h, begin, end, C.byref(s), C.byref(mx), C.byref(mn)))
        return s.value, mx.value, mn.value

    def read_paths(self, begin: int = 0, end: int | None = None) -> np.ndarray:
        end = self.n if end is None else end
        out = np.empty(max(end - begin, 1), dtype=np.float64)
        check(lib().gx_part_read_paths(self._h, begin, end, out.ctypes.data))
        return out[:end - begin]

    def kcore_init1(self, begin: int, end: int) -> None:
        check(lib().gx_part_kcore_init1(self._h, begin, end))

    def kcore_init2(self, begin: int, end: int) -> None:
        check(lib().gx_part_kcore_init2(self._h, begin, end))

    def kcore_step1(self, begin: int, end: int, k: int) -> int:
        a = C.c_uint32()
        check(lib().gx_part_kcore_step1(self._h, begin, end, k, C.byref(a)))
        return a.value

    def kcore_step2(self, begin: int, end: int) -> None:
        check(lib().gx_part_kcore_step2(self._h, begin, end))

    def kcore_check(self, begin: int, end: int) -> int:
        a = C.c_uint64()
        check(lib().gx_part_kcore_check(self._h, begin, end, C.byref(a)))
        return a.value

    def local_id(self, gid: int) -> int | None:
        if self.gid is None:
            return gid if gid < self.n else None
        hit = np.nonzero(self.gid == gid)[0]
        return int(hit[0]) if len(hit) else None


def dirgl_bfs(part: Partition, src_gid: int, delta: int = 0, max_iterations: int = 1000):
    """bfs_push.cpp BFS::go for one host; returns (dist_current of all local
    nodes, rounds, work items of the BFS rounds)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    part.bfs_init(a0, a1, INFINITY, src_gid)
    lid = part.local_id(src_gid)
    if lid is not None:
        part.bfs_first(lid, lid + 1)
    priority = 0xFFFFFFFF if delta == 0 else 0
    rounds, work = 1, 0
    while True:
        priority = (priority + delta) & 0xFFFFFFFF
        active, w = part.bfs(w0, w1, priority)
        work += w
        rounds += 1
        if not (rounds < max_iterations and active):
            break
    return part.read(DIST_CURRENT), rounds, work


def dirgl_sssp(part: Partition, src_gid: int, delta: int = 0, max_iterations: int = 100000):
    """sssp_push.cpp SSSP::go for one host; returns (dist_current of all local
    nodes, rounds, work edges of the SSSP rounds)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    part.sssp_init(a0, a1, INFINITY, src_gid)
    lid = part.local_id(src_gid)
    if lid is not None:
        part.sssp_first(lid, lid + 1)
    priority = 0xFFFFFFFF if delta == 0 else 0
    rounds, work = 1, 0
    while True:
        priority = (priority + delta) & 0xFFFFFFFF
        active, w = part.sssp(w0, w1, priority)
        work += w
        rounds += 1
        if not (rounds < max_iterations and active):
            break
    return part.read(DIST_CURRENT), rounds, work


def dirgl_cc(part: Partition, max_iterations: int = 100000):
    """cc_push.cpp ConnectedComp::go for one host; returns (comp_current of all
    local nodes, components among the masters, rounds)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    part.cc_init(a0, a1)
    part.cc_first(w0, w1)
    rounds = 1
    while True:
        active = part.cc(w0, w1)
        rounds += 1
        if not (rounds < max_iterations and active):
            break
    return part.read(COMP_CURRENT), part.cc_check(*part.range(RANGE_MASTERS)), rounds


def dirgl_pagerank(part: Partition, tolerance: float, alpha: float = PR_ALPHA, max_iterations: int = 1000):
    """pagerank_push.cpp InitializeGraph::go + PageRank::go for one host;
    returns (value of all local nodes as float32, rounds)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    part.pr_reset(a0, a1)
    part.pr_init(w0, w1, alpha)
    rounds = 0
    while True:
        part.pr_delta(w0, w1, alpha, tolerance)
        active = part.pr(w0, w1)
        rounds += 1
        if not (rounds < max_iterations and active):
            break
    return part.read_float(PR_VALUE), rounds


def dirgl_bc(part: Partition, sources):
    """bc_level.cpp's BC::go over the given source global ids for one host;
    returns the betweeness_centrality of all local nodes (float32)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    m0, m1 = part.range(RANGE_MASTERS)
    part.bc_init(a0, a1)
    for s in sources:
        part.bc_iter(a0, a1, INFINITY, int(s))
        rounds = 0
        while True:
            dga = part.bc_forward(w0, w1, rounds)
            rounds += 1
            if not dga:
                break
        if rounds > 2:
            part.bc_middle(m0, m1, INFINITY)
            for r in range(rounds - 2, 0, -1):
                part.bc_backward(w0, w1, r)
            part.bc_accumulate(m0, m1)
    return part.read_float(BC_VALUE)


def dirgl_kcore(part: Partition, k: int, max_iterations: int = 10000):
    """kcore_push.cpp InitializeGraph1::go + KCoreStep1::go for one host;
    returns (flag of all local nodes, alive masters, rounds)."""
    a0, a1 = part.range(RANGE_ALL)
    w0, w1 = part.range(RANGE_WITH_EDGES)
    part.kcore_init1(a0, a1)
    part.kcore_init2(w0, w1)
    rounds = 0
    while True:
        active = part.kcore_step1(w0, w1, k)
        part.kcore_step2(w0, w1)
        rounds += 1
        if not (rounds < max_iterations and active):
            break
    alive = part.kcore_check(*part.range(RANGE_MASTERS))
    return part.read(FLAG), alive, rounds
