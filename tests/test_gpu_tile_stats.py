"""GPU: per-tile statistics (the learned planner's features and labels).

When every (tile, fiber) group holds a single entry -- one-row tiles over
distinct coordinates -- the group pass sums the per-tile statistics itself
and the separate statistics pass is skipped. FCC_SINGLE_GROUPS=0 keeps the
separate pass. Both must record identical features and labels, and the
contraction's outputs must not change. The count features also have closed
forms at one-row tiles: log1p(entries of the row) for the entry and key
counts."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result

from oracle import oracle
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [("c1", lambda: W.c1(0.05)), ("c2", lambda: W.c2(0.05)), ("c4", lambda: W.c4(0.02))]


def run(ctx, wl, single, **kw):
    a, b = wl.tensors()
    os.environ["FCC_SINGLE_GROUPS"] = "1" if single else "0"
    try:
        r = ctx.contract(a, list(wl.a_modes), b, list(wl.b_modes), tile_hint=1, collect_stats=True, **kw)
        f, l = ctx.tile_stats()
        st = r.stats
        c, v = r.to_host()
        r.free()
    finally:
        os.environ.pop("FCC_SINGLE_GROUPS", None)
    return f, l, c, v, st


@pytest.mark.parametrize("name,mk", CASES)
def test_single_entry_groups_match_the_statistics_pass(gpu_ctx, name, mk):
    wl = mk()
    print(f"tile stats {wl.name}", flush=True)
    f1, l1, c1, v1, st1 = run(gpu_ctx, wl, True)
    f0, l0, c0, v0, st0 = run(gpu_ctx, wl, False)
    assert st1["row_tile"] == 1 and st0["row_tile"] == 1
    assert f1.shape == f0.shape and f1.shape[0] > 0
    assert np.array_equal(f1, f0), f"features differ at tiles {np.nonzero((f1 != f0).any(axis=1))[0][:10]}"
    assert np.array_equal(l1, l0)
    # outputs: the same coordinate set, values within the fp64 summation-order bound
    o1, o0 = np.lexsort(c1[::-1]), np.lexsort(c0[::-1])
    assert np.array_equal(c1[:, o1], c0[:, o0])
    assert np.allclose(v1[o1], v0[o0], rtol=1e-12, atol=1e-14)
    if wl.b is not None:
        return
    # self-contractions (every key is matched): closed forms at one-row tiles,
    # the entries and the distinct keys of each row
    a, _ = wl.tensors()
    ext = a.extents()
    xm = [m for m in range(a.order) if m not in wl.a_modes]
    row = np.zeros(a.nnz, dtype=np.int64)
    for m in xm:
        row = row * int(ext[m]) + a.coords[m].astype(np.int64)
    per_row = np.bincount(row)
    per_row = per_row[per_row > 0]
    want = np.log1p(per_row.astype(np.float32))
    assert np.allclose(np.sort(f1[:, 0]), np.sort(want), rtol=1e-6)
    assert np.allclose(np.sort(f1[:, 1]), np.sort(want), rtol=1e-6)


def test_single_entry_groups_parity(gpu_ctx):
    """One-row tiles with the fused statistics, learned planner: outputs match
    the oracle."""
    gpu_ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    wl = W.c2(0.05)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=128, high_precision=True)
    for planner in ("exact", "learned"):
        r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), tile_hint=1, planner=planner)
        assert r.stats["row_tile"] == 1
        c, v = r.to_host()
        r.free()
        assert_same_result(c, v, ref.coords, ref.vals, rel=1e-12, abs_floor=1e-14, what=f"c2 TL=1 {planner}",
                           absum=ref.absum, nterms=ref.nterms)
