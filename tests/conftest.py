import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fastcc-benchmark_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


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


def sort_order(coords):
    if coords.shape[0] == 0 or coords.shape[1] <= 1:
        return np.arange(coords.shape[1])
    return np.lexsort(coords[::-1])


def sort_result(coords, vals):
    o = sort_order(coords)
    return coords[:, o], vals[o]


PARITY_LOG = os.environ.get("FCC_PARITY_LOG")


def record_parity(what, **fields):
    """Appends one JSON line per checked result to $FCC_PARITY_LOG (when set):
    the error figures the parity tables are made from."""
    if not PARITY_LOG:
        return
    import json
    with open(PARITY_LOG, "a") as f:
        f.write(json.dumps({"what": what, **fields}) + "\n")


def assert_same_result(gc, gv, rc, rv, rel=1e-12, abs_floor=1e-14, what="", absum=None, nterms=None):
    """Sorted coordinate sets bit-exact; values |g - r| <= max(rel*|r|, abs_floor).

    When the oracle supplies, per output, the sum of |terms| and the number of
    terms, an output whose sum cancels may instead meet the worst-case error of
    an fp64 sum in an arbitrary order, (n+1) * 2^-53 * sum|terms| (one rounding
    per product and per addition); the number of outputs that needed this is
    printed, and must stay a small fraction. Non-finite values must match
    exactly (the same infinities and NaN positions). Returns the maximum
    relative error, the outputs outside the 1e-12 bound and those that met
    only the summation-order bound."""
    gc, gv = sort_result(np.asarray(gc), np.asarray(gv))
    o = sort_order(np.asarray(rc))
    rc, rv = np.asarray(rc)[:, o], np.asarray(rv)[o]
    assert gc.shape == rc.shape, f"{what}: nnz/order differ {gc.shape} vs {rc.shape}"
    assert np.array_equal(gc, rc), f"{what}: coordinate sets differ"
    fin = np.isfinite(rv)
    assert np.array_equal(np.isfinite(gv), fin), f"{what}: non-finite outputs differ"
    assert np.array_equal(np.isnan(gv), np.isnan(rv)), f"{what}: NaN outputs differ"
    assert np.array_equal(gv[np.isinf(rv)], rv[np.isinf(rv)]), f"{what}: infinite outputs differ"
    gv, rv = gv[fin], rv[fin]
    if absum is not None:
        absum = np.asarray(absum)[o][fin]
        nterms = np.asarray(nterms)[o][fin]
        o = np.arange(rv.size)
    err = np.abs(gv - rv)
    tol = np.maximum(rel * np.abs(rv), abs_floor)
    strict_bad = err > tol
    if absum is not None and nterms is not None and strict_bad.any():
        ab, nt = np.asarray(absum)[o], np.asarray(nterms)[o].astype(np.float64)
        tol_order = (nt + 1.0) * 2.0 ** -53 * ab
        bad = err > np.maximum(tol, tol_order)
        n_relaxed = int(strict_bad.sum() - bad.sum())
        print(f"{what}: {n_relaxed} of {rv.size} outputs cancel and meet the summation-order bound")
        assert n_relaxed <= max(10, rv.size // 10000), f"{what}: too many cancelling outputs ({n_relaxed})"
    else:
        bad = strict_bad
    relerr = np.abs(gv - rv) / np.maximum(np.abs(rv), 1e-300)
    worst = float(np.max(relerr)) if rv.size else 0.0
    assert not bad.any(), f"{what}: {bad.sum()} values out of tolerance; worst {worst}"
    stats = {"outputs": int(rv.size), "max_rel_err": worst,
             # where the relative bound, not the absolute floor, is the tolerance
             "max_rel_err_rel_bound": float(np.max(relerr[rel * np.abs(rv) >= abs_floor]))
             if (rel * np.abs(rv) >= abs_floor).any() else 0.0,
             "outside_1e-12": int(strict_bad.sum()), "order_bound_only": int(strict_bad.sum() - bad.sum())}
    record_parity(what, **stats)
    return stats


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """torch ships its own HIP runtime, which only initialises when it comes
    first in the process; GPU sessions start it before any library context."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass
    yield


@pytest.fixture(scope="session")
def gpu_ctx():
    from fastcc_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
