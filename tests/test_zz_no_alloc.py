"""gg_bfs_distances_noalloc, the allocation-free variant of gg_bfs_distances
(its traversal state lives in a fixed device array of the library), called
through the C ABI: two calls in a row, the second on a smaller graph, each
equal to gg_bfs_distances on the same input and to the oracle's BFS
(src/algorithms/bfs.jl:31-73); a graph above the array's limit is an
ArgumentError raised before any launch."""
import ctypes as C

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(oracle):
    import torch
    import gpugraphs as gg
    from gpugraphs import _lib
    assert torch.cuda.is_available()
    return gg, torch, _lib, oracle


def _noalloc(_lib, A, source):
    import torch
    dist = torch.empty(A.m, dtype=torch.int32, device="cuda:0")
    d = A.descriptor()
    _lib.check(_lib.lib().gg_bfs_distances_noalloc(_lib.context(0), C.byref(d), int(source),
                                                    C.c_void_p(dist.data_ptr()), 1),
               "gg_bfs_distances_noalloc")
    torch.cuda.synchronize()
    return dist.cpu().numpy()


def _existing(gg, torch, A, source):
    d = gg.bfs_distances(A, int(source))
    torch.cuda.synchronize()
    return d.cpu().numpy()


@pytest.mark.parametrize("layout", ["csr", "sell_hybrid", "sell_relabel"])
def test_noalloc_bfs_equals_bfs_distances_two_calls(env, layout):
    gg, torch, _lib, orc = env
    from gpugraphs import generators as GEN
    R = gg.ROCBackend(0)
    big = GEN.rmat(16, seed=41)
    small = GEN.rmat(11, seed=42)
    for H, tag in ((big, "first, rmat16"), (small, "second, rmat11")):
        if layout == "csr":
            A = H.to_csr(R)
        elif layout == "sell_hybrid":
            A = H.to_sell(R, 64, 4096, 64, 256)
        else:
            A = H.to_sell(R, 64, 1, 256, 1024, relabel=True, sort_entries=True)
        Ao = orc.csr(H.m, H.n, H.rowptr, H.colval, np.ones(H.nnz, np.uint8))
        deg = H.degrees()
        for s in (int(np.argmax(deg)) + 1, int(np.nonzero(deg)[0][-1]) + 1):
            print_case("noalloc bfs", layout, tag, s)
            got = _noalloc(_lib, A, s)
            np.testing.assert_array_equal(got, _existing(gg, torch, A, s))
            np.testing.assert_array_equal(got, orc.bfs_distances(Ao, s))


def test_noalloc_bfs_rejects_graphs_above_the_limit(env):
    gg, torch, _lib, orc = env
    n = (1 << 26) + 1
    rowptr = torch.ones(2, dtype=torch.int32, device="cuda:0")   # never read: rejected first
    d = _lib.GGMatrix()
    d.format, d.val_dtype = _lib.CSR, _lib.BOOL
    d.m = d.n = n
    d.nnz = 0
    d.ptr = d.idx = d.val = rowptr.data_ptr()
    d.flags = _lib.FLAG_UNIT_PATTERN
    dist = torch.empty(1, dtype=torch.int32, device="cuda:0")
    print_case("noalloc bfs above the limit", n)
    st = _lib.lib().gg_bfs_distances_noalloc(_lib.context(0), C.byref(d), 1,
                                            C.c_void_p(dist.data_ptr()), 1)
    assert st == 2   # ArgumentError
    assert b"at most" in _lib.lib().gg_last_error()
