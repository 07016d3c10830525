import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fastcc-benchmark_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


# ------------------------------------------------------------------ helpers

def rand_tensor(rng, dims, nnz, values="uniform"):
    """Unique random coordinates (numpy only; tiny sizes)."""
    from fastcc_amd.workloads import SparseTensor
    space = int(np.prod(dims))
    nnz = min(nnz, space)
    lin = rng.choice(space, size=nnz, replace=False)
    coords = np.array(np.unravel_index(lin, dims), dtype=np.uint32).reshape(len(dims), nnz)
    if values == "ones":
        vals = np.ones(nnz)
    elif values == "smallint":
        vals = rng.choice([-3.0, -2.0, -1.0, 1.0, 2.0, 3.0], size=nnz)
    elif values == "zeros":
        vals = rng.choice([0.0, 1.0, -0.0, 2.0], size=nnz)
    else:
        vals = rng.uniform(-1, 1, size=nnz)
    return SparseTensor(coords, vals)


def densify(t, dims=None):
    dims = dims or tuple(int(x) for x in t.extents())
    d = np.zeros(dims)
    if t.nnz:
        d[tuple(t.coords.astype(np.int64))] = t.vals
    return d


def sort_result(coords, vals):
    if coords.shape[0] == 0 or coords.shape[1] <= 1:
        return coords, vals
    order = np.lexsort(coords[::-1])
    return coords[:, order], vals[order]


def assert_same_result(gc, gv, rc, rv, rel=1e-12, abs_floor=1e-14, what=""):
    """Sorted coordinate sets bit-exact; values |g - r| <= max(rel*|r|, abs_floor)."""
    gc, gv = sort_result(np.asarray(gc), np.asarray(gv))
    rc, rv = sort_result(np.asarray(rc), np.asarray(rv))
    assert gc.shape == rc.shape, f"{what}: nnz/order differ {gc.shape} vs {rc.shape}"
    assert np.array_equal(gc, rc), f"{what}: coordinate sets differ"
    tol = np.maximum(rel * np.abs(rv), abs_floor)
    bad = np.abs(gv - rv) > tol
    assert not bad.any(), (f"{what}: {bad.sum()} values out of tolerance; worst "
                           f"{np.max(np.abs(gv - rv) / np.maximum(np.abs(rv), 1e-300))}")


@pytest.fixture(scope="session")
def gpu_ctx():
    from fastcc_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
