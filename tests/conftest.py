"""Shared test setup: import paths, the ``gpu`` marker, helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gpugraphs.jl_amd")
for p in (ROOT, PKG_DIR, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


def _have_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this container")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.lib()
    return orc


def print_case(*parts):
    """Name the case before a GPU step so a fault report says which one."""
    print("CASE:", *parts, flush=True)


def random_csr(rng, m, n, p, dtype=np.float32, sorted_rows=True):
    """scipy CSR with density p and U[0,1) values (like sprand, test/spmv.jl:12)."""
    import scipy.sparse as sp
    S = sp.random(m, n, density=p, format="csr", random_state=rng, dtype=np.float64)
    S.data = rng.random(S.nnz)
    S = S.astype(dtype) if dtype != np.bool_ else S.astype(np.bool_)
    if sorted_rows:
        S.sort_indices()
    return S
