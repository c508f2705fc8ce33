"""gx_cc_noalloc (connected components with the kernels' scratch in a static
device array, for callers that cannot allocate) against gx_cc, through the C
ABI: two calls in a row, the second on a smaller graph, so the second starts
from the first call's leftover scratch; and the size limit."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _labels_noalloc(gx, G):
    import torch
    from galois_amd import _lib
    out = torch.full((G.n,), -1, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().gx_cc_noalloc(G.handle, out.data_ptr()))
    return out.cpu().numpy().view(np.uint32)


def test_cc_noalloc_matches_cc(gx):
    cases = [gx.Graph.rmat(22, 16, symmetric=True, seed=11),  # n = 2^22, the largest allowed
             gx.Graph.rmat(15, 8, symmetric=True, seed=12)]   # then a smaller graph
    for G in cases:
        print(f"cc_noalloc n={G.n} m={G.m}", flush=True)
        got = _labels_noalloc(gx, G)
        exp = gx.connected_components(G)[0].cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got, exp)
        # a second call on the same graph, from this call's leftover scratch
        np.testing.assert_array_equal(_labels_noalloc(gx, G), exp)


def test_cc_noalloc_rejects_graphs_beyond_its_scratch(gx):
    import torch
    from galois_amd import _lib
    G = gx.Graph.rmat(23, 4, symmetric=True, seed=13)
    out = torch.empty(G.n, dtype=torch.int32, device="cuda")
    rc = _lib.lib().gx_cc_noalloc(G.handle, out.data_ptr())
    assert rc == _lib.GX_ERR_CAPACITY
    assert "at most 4194304" in _lib.last_error()
