"""ctypes binding of libgalois_amd.so (the C ABI declared in include/galois_amd.h).

The library is loaded from the package directory only; a missing library is a
hard error (there is no CPU fallback for any entry point).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgalois_amd.so"

GX_OK = 0
GX_ERR_INVALID = 1
GX_ERR_HIP = 2
GX_ERR_ALLOC = 3
GX_ERR_IO = 4
GX_ERR_CAPACITY = 5
GX_ERR_NO_DEVICE = 6

GX_GRAPH_SYMMETRIC = 1
GX_GRAPH_WEIGHTED = 2
GX_GRAPH_FWEIGHTED = 4

GX_DIST_INFINITY = 2147483646
GX_BC_INFINITY = 1073741823


class gx_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("launches", C.c_uint32),
        ("edges_visited", C.c_uint64),
        ("items", C.c_uint64),
        ("residual", C.c_double),
        ("elapsed_ms", C.c_double),
        ("main_kernel_ms", C.c_double),
        ("main_kernel_launches", C.c_uint32),
        ("reserved", C.c_uint32),
        ("main_kernel_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class gx_pb_layout_info(C.Structure):
    _fields_ = [(name, C.c_int64) for name in ("bins", "groups", "pieces", "tasks", "sliced_bins", "jobs_long",
                                                "jobs_mid", "jobs_short", "task_edges")] + [
        ("built", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class gx_host_csr(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("m", C.c_int64),
        ("row", C.POINTER(C.c_int32)),
        ("col", C.POINTER(C.c_int32)),
        ("edge_data", C.c_void_p),
        ("edge_size", C.c_uint32),
        ("version", C.c_uint32),
    ]


class GxError(RuntimeError):
    """A non-zero gx_status from the library, with gx_last_error's message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"gx error {code}: {message}")
        self.code = code
        self.message = message


P = C.c_void_p
I64 = C.c_int64
U32 = C.c_uint32
U64 = C.c_uint64
D = C.c_double
F = C.c_float
INT = C.c_int
PG = C.POINTER(C.c_void_p)
PST = C.POINTER(gx_stats)

# name -> (restype, argtypes); every function declared in include/galois_amd.h.
SIGNATURES: dict[str, tuple] = {
    "gx_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "gx_init": (INT, [INT]),
    "gx_set_profiling": (None, [INT]),
    "gx_tune_set": (INT, [C.c_char_p, C.c_int64]),
    "gx_tune_get": (INT, [C.c_char_p, C.POINTER(C.c_int64)]),
    "gx_tune_effective": (INT, [C.c_char_p, C.c_int64, C.c_int64, C.POINTER(C.c_int64)]),
    "gx_pagerank_layout_info": (INT, [P, C.POINTER(gx_pb_layout_info)]),
    "gx_graph_create": (INT, [I64, I64, P, P, P, U32, PG]),
    "gx_graph_create_device": (INT, [I64, I64, P, P, P, U32, PG]),
    "gx_graph_free": (INT, [P]),
    "gx_graph_info": (INT, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(U32)]),
    "gx_graph_set_stream": (INT, [P, P]),
    "gx_graph_device_arrays": (INT, [P, PG, PG, PG, PG, PG, PG]),
    "gx_graph_download": (INT, [P, P, P, P]),
    "gx_graph_download_transpose": (INT, [P, P, P, P]),
    "gx_graph_build_transpose": (INT, [P]),
    "gx_gr_read": (INT, [C.c_char_p, C.POINTER(gx_host_csr)]),
    "gx_gr_write": (INT, [C.c_char_p, C.POINTER(gx_host_csr)]),
    "gx_host_csr_free": (None, [C.POINTER(gx_host_csr)]),
    "gx_gen_rmat": (INT, [U32, U32, D, D, D, U64, INT, INT, U32, U32, PG]),
    "gx_gen_grid": (INT, [U32, U32, U32, U64, U32, U32, PG]),
    "gx_graph_union": (INT, [P, U32, PG, P]),
    "gx_pick_sources": (INT, [I64, P, U32, U32, P]),
    "gx_degree_power_law": (INT, [I64, I64, P, C.POINTER(INT)]),
    "gx_bfs": (INT, [P, U32, P, PST]),
    "gx_bfs_do": (INT, [P, U32, U32, U32, P, P, PST]),
    "gx_sssp_delta": (INT, [P, U32, U32, P, PST]),
    "gx_sssp_batched": (INT, [P, P, U32, U32, P, PST]),
    "gx_cc": (INT, [P, P, PST]),
    "gx_cc_noalloc": (INT, [P, P]),
    "gx_kcore": (INT, [P, U32, P, P, C.POINTER(U64), PST]),
    "gx_pagerank_pull": (INT, [P, D, D, U32, INT, P, PST]),
    "gx_pagerank_pull_ref": (INT, [P, INT, F, F, U32, P, PST]),
    "gx_pagerank_push": (INT, [P, F, F, U32, P, PST]),
    "gx_pagerank_weighted": (INT, [P, P, P, D, D, U32, INT, P, PST]),
    "gx_pagerank_adjoint": (INT, [P, P, P, P, D, D, U32, INT, P, P, PST]),
    "gx_bc": (INT, [P, P, U32, P, PST]),
    "gx_bc_batched": (INT, [P, P, U32, P, PST]),
    # D-IrGL personality partition
    "gx_part_create": (INT, [INT, PG]),
    "gx_part_free": (INT, [P]),
    "gx_part_load": (INT, [P, U64, U64, P, P, P, P, U32, U32, U32]),
    "gx_part_range": (INT, [P, INT, C.POINTER(U32), C.POINTER(U32)]),
    "gx_part_get": (INT, [P, INT, U32, C.POINTER(U32)]),
    "gx_part_set": (INT, [P, INT, U32, U32]),
    "gx_part_add": (INT, [P, INT, U32, U32]),
    "gx_part_min": (INT, [P, INT, U32, U32, C.POINTER(INT)]),
    "gx_part_fill": (INT, [P, INT, U32, U32, U32]),
    "gx_part_read": (INT, [P, INT, U32, U32, P]),
    "gx_part_bitset_get": (INT, [P, INT, P]),
    "gx_part_bitset_reset": (INT, [P, INT, U32, U32]),
    "gx_part_bfs_init": (INT, [P, U32, U32, U32, U64]),
    "gx_part_bfs_first": (INT, [P, U32, U32]),
    "gx_part_bfs": (INT, [P, U32, U32, U32, C.POINTER(U32), C.POINTER(U32)]),
    "gx_part_bfs_check": (INT, [P, U32, U32, U32, C.POINTER(U64), C.POINTER(U32)]),
    "gx_part_sssp_init": (INT, [P, U32, U32, U32, U64]),
    "gx_part_sssp_first": (INT, [P, U32, U32]),
    "gx_part_sssp": (INT, [P, U32, U32, U32, C.POINTER(U32), C.POINTER(U32)]),
    "gx_part_sssp_check": (INT, [P, U32, U32, U32, C.POINTER(U64), C.POINTER(U64), C.POINTER(U32)]),
    "gx_part_cc_init": (INT, [P, U32, U32]),
    "gx_part_cc_first": (INT, [P, U32, U32]),
    "gx_part_cc": (INT, [P, U32, U32, C.POINTER(U32)]),
    "gx_part_cc_check": (INT, [P, U32, U32, C.POINTER(U64)]),
    "gx_part_pr_reset": (INT, [P, U32, U32]),
    "gx_part_pr_init": (INT, [P, U32, U32, F]),
    "gx_part_pr_delta": (INT, [P, U32, U32, F, F]),
    "gx_part_pr": (INT, [P, U32, U32, C.POINTER(U32)]),
    "gx_part_pr_check": (INT, [P, U32, U32, F, P]),
    "gx_part_bc_init": (INT, [P, U32, U32]),
    "gx_part_bc_iter": (INT, [P, U32, U32, U32, U64]),
    "gx_part_bc_forward": (INT, [P, U32, U32, U32, C.POINTER(U32)]),
    "gx_part_bc_middle": (INT, [P, U32, U32, U32]),
    "gx_part_bc_backward": (INT, [P, U32, U32, U32]),
    "gx_part_bc_accumulate": (INT, [P, U32, U32]),
    "gx_part_bc_check": (INT, [P, U32, U32, C.POINTER(F), C.POINTER(F), C.POINTER(F)]),
    "gx_part_read_paths": (INT, [P, U32, U32, P]),
    "gx_part_kcore_init1": (INT, [P, U32, U32]),
    "gx_part_kcore_init2": (INT, [P, U32, U32]),
    "gx_part_kcore_step1": (INT, [P, U32, U32, U32, C.POINTER(U32)]),
    "gx_part_kcore_step2": (INT, [P, U32, U32]),
    "gx_part_kcore_check": (INT, [P, U32, U32, C.POINTER(U64)]),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load libgalois_amd.so (once) and attach the C signatures."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no fallback implementation)")
        # torch ships its own HIP runtime under the same soname as ROCm's.
        # Loading torch first makes this library bind to that one copy; if
        # ROCm's copy were loaded first, torch would later bind to it and
        # fail to see the GPU.
        import torch  # noqa: F401
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(4096)
    lib().gx_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != GX_OK:
        raise GxError(status, last_error())
