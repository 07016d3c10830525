"""GPU parity of batch plans: fcc_contract_batch's fused launch planned once
for its pairs' coordinates, then rerun on new values of every pair, eagerly
or as one replayed hipGraph. Every execution of every pair must equal the
oracle's contraction of that pair with the new values (materialised: the
coordinate set bit-exact and values within max(1e-12 |ref|, 1e-14); count +
checksum: count and coordinate hash exact, sums within 1e-9). Graph replays
are interleaved with eager executions, other contractions and
synchronisation, and each replay is checked."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FCC_ACC_AUTO, FCC_ACC_GHASH, FCC_ACC_LHASH, FastccError
from fastcc_amd import workloads as W
from fastcc_amd.workloads import SparseTensor

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14


def _load_mlp(ctx):
    ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))


def _bbox(ts):
    return np.max(np.stack([t.extents() for t in ts]), axis=0)


def _new_values(rng, pairs, it):
    va = [rng.uniform(-1, 1, a.nnz) for a, _ in pairs]
    vb = [rng.uniform(-1, 1, b.nnz) for _, b in pairs]
    if it == 2 and len(va) > 1:  # exact zeros: the structural support is unchanged
        va[1][::2] = 0.0
        vb[0][::3] = -0.0
    return va, vb


def _with(t, v):
    return SparseTensor(t.coords, np.ascontiguousarray(v, dtype=np.float64))


SPECS = [
    ((20, 6, 9), 300, (6, 15), 40),
    ((3, 2, 2), 5, (2, 4), 3),
    ((40, 12, 30), 4000, (12, 60), 500),
    ((1, 1, 1), 1, (1, 1), 1),
    ((30, 4, 25), 2000, (4, 70), 200),
]


@pytest.mark.parametrize("planner", ["exact", "learned"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_batch_plan_materialised(gpu_ctx, planner, force):
    print(f"batch plan materialised planner={planner} force={force}", flush=True)
    _load_mlp(gpu_ctx)
    rng = np.random.default_rng(5100)
    pairs = [(rand_tensor(rng, da, na, "uniform"), rand_tensor(rng, db, nb, "uniform")) for da, na, db, nb in SPECS]
    plan = gpu_ctx.plan_batch(pairs, [1], [0], planner=planner, force_acc=force)
    for it, graph in enumerate([False, True, True, False, True]):
        va, vb = _new_values(rng, pairs, it)
        rs = plan.execute(np.concatenate(va), np.concatenate(vb), graph=graph)
        assert len(rs) == len(pairs)
        for i, ((a, b), r) in enumerate(zip(pairs, rs)):
            ref = oracle.contract_tiled(_with(a, va[i]), (1,), _with(b, vb[i]), (0,), tile_size=8)
            c, v = r.to_host()
            assert r.order == 3
            assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS,
                               what=f"execution {it} graph={graph} pair {i}", absum=ref.absum, nterms=ref.nterms)
        # other work between replays
        x = gpu_ctx.contract(pairs[2][0], [1], pairs[2][1], [0])
        x.free()
    plan.close()


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_batch_plan_sink_c5_shape(gpu_ctx, planner):
    """Scaled pairs of the order-4 x order-3 uneven batch shape, count +
    checksum per pair, replayed as one graph with new values each time."""
    print(f"batch plan sink c5 planner={planner}", flush=True)
    _load_mlp(gpu_ctx)
    wls = W.c5_pairs(n_pairs=24, scale=0.02)
    pairs = [w.tensors() for w in wls]
    ea, eb = _bbox([p[0] for p in pairs]), _bbox([p[1] for p in pairs])
    plan = gpu_ctx.plan_batch(pairs, list(wls[0].a_modes), list(wls[0].b_modes), planner=planner, sink=True)
    rng = np.random.default_rng(5200)
    for it, graph in enumerate([True, False, True, True]):
        va, vb = _new_values(rng, pairs, it)
        rs = plan.execute(np.concatenate(va), np.concatenate(vb), graph=graph)
        for i, ((a, b), r) in enumerate(zip(pairs, rs)):
            ref = oracle.contract_tiled(_with(a, va[i]), wls[i].a_modes, _with(b, vb[i]), wls[i].b_modes,
                                        tile_size=128, sink=True, a_extents=ea, b_extents=eb)
            assert r.nnz == ref.nnz, (it, i)
            assert r.coord_hash == ref.coord_hash, (it, i)
            assert r.val_sum == pytest.approx(ref.val_sum, rel=1e-9, abs=1e-9), (it, i)
            assert r.val_wsum == pytest.approx(ref.val_wsum, rel=1e-9, abs=1e-9), (it, i)
            assert 0.0 < r.stats["ms_done"] <= r.stats["ms_total"] + 1e-6
    plan.close()


def test_batch_plan_device_values_and_errors(gpu_ctx):
    import torch
    rng = np.random.default_rng(5300)
    pairs = [(rand_tensor(rng, da, na, "smallint"), rand_tensor(rng, db, nb, "smallint")) for da, na, db, nb in SPECS]
    plan = gpu_ctx.plan_batch(pairs, [1], [0])
    va, vb = _new_values(rng, pairs, 0)
    da = torch.tensor(np.concatenate(va), device="cuda")
    db = torch.tensor(np.concatenate(vb), device="cuda")
    for graph in (False, True):
        rs = plan.execute(da, db, graph=graph)
        for i, ((a, b), r) in enumerate(zip(pairs, rs)):
            ref = oracle.contract_tiled(_with(a, va[i]), (1,), _with(b, vb[i]), (0,), tile_size=8)
            c, v = r.to_host()
            assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"device values pair {i}",
                               absum=ref.absum, nterms=ref.nterms)
    # a batch plan is not a single-contraction plan
    with pytest.raises(FastccError) as ei:
        super(type(plan), plan).execute(np.concatenate(va), np.concatenate(vb))
    assert ei.value.code == -1
    plan.close()
    # pairs that do not fit one fused launch
    big = []
    for i in range(4100):
        ca = np.array([[1023, 0], [0, 1], [1023, 5]], dtype=np.uint32)
        cb = np.array([[0, 1], [3, 4]], dtype=np.uint32)
        big.append((SparseTensor(ca, np.ones(2)), SparseTensor(cb, np.ones(2))))
    with pytest.raises(FastccError) as ei:
        gpu_ctx.plan_batch(big, [1], [0])
    assert ei.value.code == -1
