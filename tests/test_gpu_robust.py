"""Edge cases of the contraction's arithmetic and limits, each against the
oracle: non-finite input values next to the dense-key GEMM, a forced
LDS-hash accumulator over its key limit, external spans that overflow 64-bit
products, fused batches whose pair index would overflow one launch, and a
device-memory budget that makes the contraction fail cleanly."""
import numpy as np
import pytest

from conftest import assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FCC_ACC_LHASH, FastccError
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14


def host(ctx, a, am, b, bm, **kw):
    r = ctx.contract(a, list(am), b, list(bm), **kw)
    st = r.stats
    c, v = r.to_host()
    r.free()
    return c, v, st


@pytest.mark.parametrize("bad", [np.inf, -np.inf, np.nan])
@pytest.mark.parametrize("self_contr", [True, False])
def test_nonfinite_values_with_dense_keys(gpu_ctx, bad, self_contr):
    """An infinite or NaN value reaches only the outputs its products reach:
    the dense-key GEMM (zero-padded panels) must not spread it. Keys whose
    fibers are dense on both sides would be taken by the GEMM (threshold set
    so that they are) -- with a non-finite value present they stay in the
    tiles."""
    print(f"nonfinite bad={bad} self={self_contr}", flush=True)
    rng = np.random.default_rng(77)
    t = rand_tensor(rng, (6, 40, 30), 3000, "uniform")
    k = int(np.argmax(t.coords[0] == 0))  # an entry on key 0 (a dense key)
    t.vals[k] = bad
    if self_contr:
        a = b = t
        am = bm = (0,)
    else:
        a = t
        b = rand_tensor(rng, (6, 50), 200, "uniform")
        am, bm = (0,), (0,)
    ref = oracle.contract_bruteforce(a, am, b, bm)
    c, v, st = host(gpu_ctx, a, am, b, bm, dense_keys=1e-6)
    assert st["dense_keys"] == 0
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"nonfinite {bad}",
                       absum=ref.absum, nterms=ref.nterms)
    assert (~np.isfinite(v)).sum() == (~np.isfinite(ref.vals)).sum() > 0
    # without the bad value the same keys do go through the GEMM
    t.vals[k] = 1.0
    _, _, st = host(gpu_ctx, a, am, b, bm, dense_keys=1e-6)
    assert st["dense_keys"] > 0


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_forced_lhash_over_its_limit(gpu_ctx, planner):
    """Forced LDS-hash units holding more than the table's 3072-key limit seal
    the table and spill to the global hash, which must be sized for them.
    B's columns crowd into [0, 100) except one far column, so the planner's
    even column windows put 6400 outputs of each 64-row tile into its first
    window; 20 such tiles spill about 66k keys."""
    import os
    from conftest import ROOT
    from fastcc_amd.workloads import SparseTensor
    gpu_ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    rng = np.random.default_rng(31)
    rows, keys = 1280, 50
    ca = np.stack(np.meshgrid(np.arange(rows), np.arange(keys), indexing="ij")).reshape(2, -1).astype(np.uint32)
    a = SparseTensor(ca, rng.uniform(-1, 1, ca.shape[1]))
    cb = np.stack(np.meshgrid(np.arange(keys), np.arange(100), indexing="ij")).reshape(2, -1)
    cb = np.concatenate([cb, [[0], [99999]]], axis=1).astype(np.uint32)
    b = SparseTensor(cb, rng.uniform(-1, 1, cb.shape[1]))
    ref = oracle.contract_tiled(a, (1,), b, (0,), tile_size=64, high_precision=True)
    c, v, st = host(gpu_ctx, a, (1,), b, (0,), force_acc=FCC_ACC_LHASH, tile_hint=64, planner=planner)
    assert st["units_lhash"] > 0 and st["ghash_entries"] > 3 * 65536 // 4, st
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"forced lhash {planner}",
                       absum=ref.absum, nterms=ref.nterms)


def test_span_product_overflow_is_invalid(gpu_ctx):
    """Four external modes of extent 65536 span 2^64: rejected, not wrapped to 0."""
    from fastcc_amd.workloads import SparseTensor
    rng = np.random.default_rng(5)
    c = rng.integers(0, 100, size=(5, 50)).astype(np.uint32)
    c[0] %= 4
    a = SparseTensor(c, rng.uniform(-1, 1, 50))
    b = rand_tensor(rng, (4, 3), 6)
    a.given_extents = (4, 65536, 65536, 65536, 65536)
    with pytest.raises(FastccError) as ei:
        gpu_ctx.contract(a, [0], b, [0])
    assert ei.value.code == -1
    a.given_extents = (4, 256, 256, 256, 100)  # 1.7e9 rows: fits
    r = gpu_ctx.contract(a, [0], b, [0])
    assert r.nnz > 0
    r.free()


def test_batch_beyond_one_launch(gpu_ctx):
    """4100 pairs whose left span is 2^20 each: a fused launch holds at most
    4095 of them; every pair must still match the oracle."""
    from fastcc_amd.workloads import SparseTensor
    rng = np.random.default_rng(9)
    pairs, refs = [], []
    for i in range(4100):
        ca = np.stack([rng.integers(0, 1024, 3), rng.integers(0, 4, 3), rng.integers(0, 1024, 3)]).astype(np.uint32)
        ca[0, 0], ca[2, 0] = 1023, 1023  # every pair spans the full 1024 x 1024 rows
        a = SparseTensor(ca, rng.uniform(-1, 1, 3))
        cb = np.stack([rng.permutation(4)[:3], rng.integers(0, 8, 3)]).astype(np.uint32)
        b = SparseTensor(cb, rng.uniform(-1, 1, 3))
        pairs.append((a, b))
    rs = gpu_ctx.contract_batch(pairs, [1], [0])
    for i in list(range(0, 4100, 97)) + [4094, 4095, 4099]:
        a, b = pairs[i]
        ref = oracle.contract_bruteforce(a, (1,), b, (0,))
        c, v = rs[i].to_host()
        assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"pair {i}")
    assert sum(r.nnz for r in rs) == sum(oracle.contract_bruteforce(a, (1,), b, (0,)).nnz for a, b in pairs)
    for r in rs:
        r.free()


def test_memory_budget_fails_cleanly(gpu_ctx):
    """A budget too small for the contraction gives FCC_ERR_OOM, the context
    holds no more in-use memory afterwards than before, and the same call
    succeeds once the budget is lifted."""
    from fastcc_amd import Context
    ctx = Context(0)
    wl = W.c2(0.1)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=128, high_precision=True)
    before = ctx.memory_stats()
    ctx.set_memory_budget(8 << 20)
    for _ in range(2):
        with pytest.raises(FastccError) as ei:
            ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes))
        assert ei.value.code == -2
        after = ctx.memory_stats()
        assert after["in_use"] == before["in_use"]
        assert after["held"] <= 8 << 20
    ctx.set_memory_budget(0)
    c, v, _ = host(ctx, a, wl.a_modes, a, wl.b_modes)
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what="after budget",
                       absum=ref.absum, nterms=ref.nterms)
    assert ctx.memory_stats()["in_use"] == before["in_use"]
    ctx.close()
