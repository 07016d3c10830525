"""fcc_contract_noalloc through the C ABI: the contraction of fcc_contract
with every device temporary in the library's fixed scratch array and the
outputs in caller buffers. Checked against fcc_contract on the same inputs
(coordinate sets bit-exact, values within max(1e-12 |ref|, 1e-14)) on two
calls in a row, the second smaller, for both planners and both sinks; the
context's device pool does not grow during these calls."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_same_result, rand_tensor

from fastcc_amd import _abi, _as_fcc_tensor
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14


def noalloc(ctx, a, am, b, bm, capacity, **opts):
    import torch
    keep = []
    ta = _as_fcc_tensor(a, keep)
    tb = ta if b is a else _as_fcc_tensor(b, keep)
    r = len(am)
    amc = (C.c_int32 * max(r, 1))(*am)
    bmc = (C.c_int32 * max(r, 1))(*bm)
    o = ctx.options(**opts)
    d = a.order + b.order - 2 * r
    coords = torch.zeros((max(d, 1), max(capacity, 1)), dtype=torch.int32, device="cuda:0")
    vals = torch.zeros(max(capacity, 1), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    raw = _abi.FccResult()
    rc = ctx.lib.fcc_contract_noalloc(ctx.h, C.byref(ta), amc, C.byref(tb), bmc, r, C.byref(o), coords.data_ptr(),
                                      vals.data_ptr(), int(capacity), C.byref(raw))
    n = raw.nnz
    if rc != 0 or opts.get("sink"):
        return rc, n, None, None, raw
    assert raw.coord_stride == capacity
    c = coords[:d, :n].cpu().numpy().view(np.uint32) if d else np.zeros((0, n), np.uint32)
    v = vals[:n].cpu().numpy()
    return rc, n, c, v, raw


def reference(ctx, a, am, b, bm, **opts):
    r = ctx.contract(a, list(am), b, list(bm), **opts)
    c, v = r.to_host()
    out = (r.nnz, c, v, r.coord_hash, r.val_sum)
    r.free()
    return out


@pytest.fixture(scope="module")
def ctx():
    import os
    from conftest import ROOT
    from fastcc_amd import Context
    c = Context(0)
    c.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    yield c
    c.close()


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_noalloc_matches_contract_twice(ctx, planner):
    rng = np.random.default_rng(606)
    big = W.c1(0.2).tensors()
    small = (rand_tensor(rng, (30, 8, 20), 1500), rand_tensor(rng, (8, 40), 200))
    cases = [(big[0], (1, 3), big[1], (1, 2)), (small[0], (1,), small[1], (0,))]
    for a, am, b, bm in cases:  # the second call is smaller
        n, rc_, rv, _, _ = reference(ctx, a, am, b, bm, planner=planner)
        held = ctx.memory_stats()["held"]
        rc, m, c, v, raw = noalloc(ctx, a, am, b, bm, n + 100, planner=planner)
        assert rc == 0, ctx.lib.fcc_last_error(ctx.h)
        assert m == n
        assert ctx.memory_stats()["held"] == held  # nothing taken from the device pool
        assert_same_result(c, v, rc_, rv, rel=REL, abs_floor=ABS, what=f"noalloc {planner} nnz {n}")


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_noalloc_self_contraction_and_sink(ctx, planner):
    a, _ = W.c2(0.1).tensors()
    am = (0, 1)
    n, rc_, rv, h, s = reference(ctx, a, am, a, am, planner=planner, dense_keys=-1.0, summarize=True)
    rc, m, c, v, raw = noalloc(ctx, a, am, a, am, n, planner=planner)
    assert rc == 0 and m == n and raw.stats.mirrored == 1
    assert_same_result(c, v, rc_, rv, rel=REL, abs_floor=ABS, what=f"noalloc self {planner}")
    rc, m, _, _, raw = noalloc(ctx, a, am, a, am, 0, planner=planner, sink=True)
    assert rc == 0 and m == n and raw.coord_hash == h
    assert raw.val_sum == pytest.approx(s, rel=1e-10, abs=1e-10)


def test_noalloc_limits(ctx):
    rng = np.random.default_rng(7)
    a, b = rand_tensor(rng, (30, 8, 20), 1500), rand_tensor(rng, (8, 40), 200)
    n = reference(ctx, a, (1,), b, (0,))[0]
    rc, m, _, _, _ = noalloc(ctx, a, (1,), b, (0,), n - 1)
    assert rc == _abi.FCC_ERR_OVERFLOW and m == n  # the needed count comes back
    big = W.c2(1.0).tensors()[0]  # its temporaries exceed the 64 MiB scratch
    rc, _, _, _, _ = noalloc(ctx, big, (0, 1), big, (0, 1), 1 << 20)
    assert rc == _abi.FCC_ERR_INVALID
    assert b"scratch" in ctx.lib.fcc_last_error(ctx.h)
    # the entry point keeps working after a rejected call
    rc, m, _, _, _ = noalloc(ctx, a, (1,), b, (0,), n)
    assert rc == 0 and m == n
