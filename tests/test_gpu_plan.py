"""GPU parity of plans: one coordinate pattern, many value sets.

A plan keeps the index, tiles, accumulator kinds and per-unit output offsets
of a contraction; fcc_plan_execute reruns only the numeric pass on new values,
eagerly or as a replayed hipGraph. Every execution must equal the oracle's
contraction of the same coordinates with the new values. Graph replays are
interleaved with eager work and synchronisation, and each is checked."""
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


def _with_vals(t, v):
    return SparseTensor(t.coords, np.ascontiguousarray(v, dtype=np.float64))


def _check(r, a, am, b, bm, what, **kw):
    ref = oracle.contract_tiled(a, am, b, bm, tile_size=64, **kw)
    c, v = r.to_host()
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=what, absum=ref.absum,
                       nterms=ref.nterms)
    return ref


@pytest.mark.parametrize("planner", ["exact", "learned"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_plan_new_values(gpu_ctx, planner, force):
    print(f"planner={planner} force={force}", flush=True)
    _load_mlp(gpu_ctx)
    rng = np.random.default_rng(4000)
    a = rand_tensor(rng, (60, 12, 40), 6000, "uniform")
    b = rand_tensor(rng, (12, 70), 500, "uniform")
    plan = gpu_ctx.plan(a, [1], b, [0], planner=planner, force_acc=force)
    info = plan.info()
    for it in range(3):
        va, vb = rng.uniform(-1, 1, a.nnz), rng.uniform(-1, 1, b.nnz)
        if it == 2:  # exact zeros in later values: structural support is unchanged
            va[::3] = 0.0
            vb[::5] = -0.0
        r = plan.execute(va, vb)
        ref = _check(r, _with_vals(a, va), (1,), _with_vals(b, vb), (0,), f"{planner} iteration {it}")
        assert r.nnz == info["out_nnz"] == ref.nnz
    assert info["products"] == ref.products
    plan.close()


def test_plan_query_bounds(gpu_ctx):
    _load_mlp(gpu_ctx)
    wl = W.c1(0.25)
    a, b = wl.tensors()
    for planner in ("exact", "learned"):
        plan = gpu_ctx.plan(a, list(wl.a_modes), b, list(wl.b_modes), planner=planner)
        info = plan.info()
        nnz, prod, pred = plan.tiles()
        print(planner, info, flush=True)
        assert int(nnz.sum()) == info["out_nnz"]
        assert int(prod.sum()) == info["products"]
        assert info["nnz_lower_bound"] <= info["out_nnz"] <= info["nnz_upper_bound"] <= info["products"]
        assert np.all(nnz <= prod)
        if planner == "exact":
            assert np.all(pred == -1.0)
        else:
            assert np.all(pred >= 0.0)
        plan.close()


def test_plan_tiles_learned_sink_and_mirrored(gpu_ctx):
    """Per-tile output counts do not depend on the planner, the output sink or
    mirroring: a learned count + checksum plan takes them from its units'
    outputs; a mirrored self-contraction plan (whose tiles write only part of
    their bands) from a band count pass. All equal the exact planner's count
    pass on the unmirrored materialised contraction, tile by tile."""
    _load_mlp(gpu_ctx)
    for wl, self_c in ((W.c1(0.25), False), (W.c2(0.1), True), (W.c4(0.02), True)):
        a, b = wl.tensors()
        bb = a if self_c else b
        per = {}
        for planner, sink, mirror in (("exact", False, False), ("learned", True, False), ("exact", False, True),
                                      ("learned", True, True), ("learned", False, True)):
            if mirror and not self_c:
                continue
            print(wl.name, planner, sink, mirror, flush=True)
            plan = gpu_ctx.plan(a, list(wl.a_modes), bb, list(wl.b_modes), planner=planner, sink=sink, mirror=mirror,
                                tile_hint=16)
            info = plan.info()
            nnz, prod, _ = plan.tiles()
            per[(planner, sink, mirror)] = (nnz, prod, info)
            plan.close()
        nnz0, prod0, info0 = per[("exact", False, False)]
        assert int(nnz0.sum()) == info0["out_nnz"]
        for key, (nnz, prod, info) in per.items():
            assert np.array_equal(nnz, nnz0), (wl.name, key)
            assert np.array_equal(prod, prod0), (wl.name, key)


@pytest.mark.parametrize("sink", [False, True])
def test_plan_graph_replays(gpu_ctx, sink):
    """Captured once, replayed with new values each time; eager executions and
    synchronisation between replays; every replay checked."""
    import torch
    print(f"sink={sink}", flush=True)
    rng = np.random.default_rng(4100)
    a = rand_tensor(rng, (30, 16, 50), 8000, "uniform")
    b = rand_tensor(rng, (16, 90), 900, "uniform")
    plan = gpu_ctx.plan(a, [1], b, [0], sink=sink)
    dev = torch.device("cuda:0")
    for it in range(4):
        va = torch.from_numpy(rng.uniform(-1, 1, a.nnz)).to(dev)
        vb = torch.from_numpy(rng.uniform(-1, 1, b.nnz)).to(dev)
        torch.cuda.synchronize()
        rg = plan.execute(va, vb, graph=True)
        ha, hb = va.cpu().numpy(), vb.cpu().numpy()
        ref = oracle.contract_tiled(_with_vals(a, ha), (1,), _with_vals(b, hb), (0,), tile_size=64, sink=sink)
        if sink:
            assert rg.nnz == ref.nnz and rg.coord_hash == ref.coord_hash
            assert abs(rg.val_sum - ref.val_sum) <= 1e-9 * max(1.0, abs(ref.val_sum))
        else:
            c, v = rg.to_host()
            assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"replay {it}",
                               absum=ref.absum, nterms=ref.nterms)
        # eager execution of the same values in between replays
        re = plan.execute(va, vb, graph=False)
        if sink:
            assert re.nnz == rg.nnz and re.coord_hash == rg.coord_hash
        else:
            c2, v2 = re.to_host()
            assert_same_result(c2, v2, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"eager {it}",
                               absum=ref.absum, nterms=ref.nterms)
        torch.cuda.synchronize()
    plan.close()


def test_plan_self_contraction(gpu_ctx):
    """Self-contraction plan (a and b the same tensor): one value array."""
    wl = W.c2(0.1)
    a, _ = wl.tensors()
    plan = gpu_ctx.plan(a, list(wl.a_modes), a, list(wl.b_modes))
    rng = np.random.default_rng(4200)
    for graph in (False, True, True):
        va = rng.lognormal(0, 1, a.nnz)
        r = plan.execute(va, None, graph=graph)
        t = _with_vals(a, va)
        _check(r, t, wl.a_modes, t, wl.b_modes, f"self graph={graph}")
    plan.close()


def test_plan_batch_modes(gpu_ctx):
    rng = np.random.default_rng(4300)
    a = rand_tensor(rng, (6, 40, 20, 30), 9000, "uniform")
    b = rand_tensor(rng, (20, 6, 50), 2500, "uniform")
    plan = gpu_ctx.plan(a, [2], b, [0], a_batch=[0], b_batch=[1])
    for graph in (False, True):
        va, vb = rng.uniform(-1, 1, a.nnz), rng.uniform(-1, 1, b.nnz)
        r = plan.execute(va, vb, graph=graph)
        ref = oracle.contract_tiled(_with_vals(a, va), (2,), _with_vals(b, vb), (0,), a_batch=(0,), b_batch=(1,),
                                    tile_size=32)
        c, v = r.to_host()
        assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"batch plan graph={graph}",
                           absum=ref.absum, nterms=ref.nterms)
    plan.close()


def test_plan_errors(gpu_ctx):
    rng = np.random.default_rng(4400)
    a = rand_tensor(rng, (5, 6), 10, "uniform")
    b = rand_tensor(rng, (6, 4), 8, "uniform")
    plan = gpu_ctx.plan(a, [1], b, [0])
    with pytest.raises(FastccError):
        plan.execute(None, np.ones(b.nnz))
    plan.close()
    e = SparseTensor(np.zeros((2, 0), np.uint32), np.zeros(0))
    plan = gpu_ctx.plan(e, [1], b, [0])
    assert plan.execute(np.zeros(0), np.ones(b.nnz)).nnz == 0
    plan.close()


def test_plan_nonfinite_values_with_dense_keys(gpu_ctx):
    """A plan that multiplies dense keys on the matrix cores refuses values
    holding inf or NaN (a zero-padded panel would spread them to cells they
    never reach) and keeps working with finite values; a plan made without
    dense keys takes them, and matches the oracle."""
    wl = W.c2(0.1)
    a, _ = wl.tensors()
    rng = np.random.default_rng(4700)
    r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), dense_keys=1e-3)
    assert r.stats["dense_keys"] > 0
    r.free()
    for dk in (1e-3, -1.0):
        plan = gpu_ctx.plan(a, list(wl.a_modes), a, list(wl.b_modes), dense_keys=dk)
        v = rng.uniform(-1, 1, a.nnz)
        v[7] = np.inf
        if dk > 0:
            with pytest.raises(FastccError) as ei:
                plan.execute(v)
            assert ei.value.code == -1
        else:
            r = plan.execute(v, graph=True)
            ref = oracle.contract_tiled(_with_vals(a, v), wl.a_modes, _with_vals(a, v), wl.b_modes, tile_size=128)
            c, vv = r.to_host()
            assert_same_result(c, vv, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what="inf without dense keys",
                               absum=ref.absum, nterms=ref.nterms)
        v[7] = 0.5
        for graph in (False, True):
            r = plan.execute(v, graph=graph)
            ref = oracle.contract_tiled(_with_vals(a, v), wl.a_modes, _with_vals(a, v), wl.b_modes, tile_size=128)
            c, vv = r.to_host()
            assert_same_result(c, vv, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"dense_keys {dk}",
                               absum=ref.absum, nterms=ref.nterms)
        plan.close()
