"""GPU parity of batch modes and of fused batches of independent contractions.

Batch modes follow Tensor::multiply(other, left_contr, left_batch, right_contr,
right_batch) (contract.hpp:285-367): a batch pair must agree like a
contraction pair and leads the output coordinate (index_tensors.hpp:318-353).
The oracle's batch form is pinned against numpy einsum in test_oracle.py.

fcc_contract_batch runs n pairs as one batch-mode contraction over a pair
index; each pair's result must equal the oracle's contraction of that pair,
with the summaries linearised over the per-mode maximum extents of all pairs."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_GHASH, FCC_ACC_LHASH, FastccError
from fastcc_amd import workloads as W
from fastcc_amd.workloads import SparseTensor

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14

BATCH_CASES = [
    # (dims A, nnz A, a_modes, a_batch, dims B, nnz B, b_modes, b_batch)
    ((4, 5, 6), 60, (2,), (0,), (4, 6, 3), 40, (1,), (0,)),
    ((3, 5, 4, 6), 90, (1, 3), (2,), (6, 4, 5, 2), 70, (3, 0), (1,)),
    ((3, 4, 5), 40, (), (0, 2), (5, 3, 6), 30, (), (0, 1)),
    ((2, 3, 4), 20, (1,), (2, 0), (3, 2, 4), 22, (0,), (2, 1)),
    ((12, 40, 30), 3000, (1,), (0,), (12, 40, 50), 2000, (1,), (0,)),
]


def _load_mlp(ctx):
    ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))


@pytest.mark.parametrize("case", range(len(BATCH_CASES)))
@pytest.mark.parametrize("values", ["uniform", "smallint", "zeros"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_batch_modes(gpu_ctx, case, values, force):
    print(f"case={case} values={values} force={force}", flush=True)
    rng = np.random.default_rng(3000 + case)
    da, na, am, ab, db, nb, bm, bb = BATCH_CASES[case]
    a = rand_tensor(rng, da, na, values)
    b = rand_tensor(rng, db, nb, values)
    ref = oracle.contract_bruteforce(a, am, b, bm, a_batch=ab, b_batch=bb)
    r = gpu_ctx.contract(a, list(am), b, list(bm), a_batch=list(ab), b_batch=list(bb), force_acc=force)
    c, v = r.to_host()
    assert r.order == ref.order
    assert r.stats["products"] == ref.products
    exact = values != "uniform"
    assert_same_result(c, v, ref.coords, ref.vals, rel=0 if exact else REL, abs_floor=0 if exact else ABS,
                       what=f"batch case {case}", absum=None if exact else ref.absum,
                       nterms=None if exact else ref.nterms)
    r.free()


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_batch_modes_midsize(gpu_ctx, planner):
    """Order-4 batch contraction large enough for many tiles of every kind."""
    print(f"planner={planner}", flush=True)
    _load_mlp(gpu_ctx)
    rng = np.random.default_rng(3100)
    a = rand_tensor(rng, (8, 300, 40, 50), 60000, "uniform")
    b = rand_tensor(rng, (40, 8, 200), 8000, "uniform")
    ref = oracle.contract_tiled(a, (2,), b, (0,), a_batch=(0,), b_batch=(1,), tile_size=64)
    r = gpu_ctx.contract(a, [2], b, [0], a_batch=[0], b_batch=[1], planner=planner)
    c, v = r.to_host()
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what="midsize batch", absum=ref.absum,
                       nterms=ref.nterms)
    r.free()


def test_batch_mode_errors(gpu_ctx):
    rng = np.random.default_rng(3200)
    a = rand_tensor(rng, (3, 4, 5), 20, "uniform")
    b = rand_tensor(rng, (3, 4, 6), 20, "uniform")
    with pytest.raises(FastccError):  # a mode both batch and contracted
        gpu_ctx.contract(a, [0], b, [0], a_batch=[0], b_batch=[1])
    with pytest.raises(FastccError):  # lists of different length
        gpu_ctx.contract(a, [1], b, [1], a_batch=[0], b_batch=[])


def _pairs(seed, specs):
    """Uneven pairs of order-3 A (x, k, y) and order-2 B (k, z)."""
    rng = np.random.default_rng(seed)
    out = []
    for (dims_a, na, dims_b, nb, values) in specs:
        out.append((rand_tensor(rng, dims_a, na, values), rand_tensor(rng, dims_b, nb, values)))
    return out


def _bbox(ts):
    return np.max(np.stack([t.extents() if t.nnz else np.zeros(t.order, np.uint64) for t in ts]), axis=0)


PAIR_SPECS = [
    ((20, 6, 9), 300, (6, 15), 40, "uniform"),
    ((3, 2, 2), 5, (2, 4), 3, "uniform"),
    ((40, 12, 30), 4000, (12, 60), 500, "uniform"),
    ((7, 5, 3), 60, (5, 2), 8, "zeros"),
    ((1, 1, 1), 1, (1, 1), 1, "uniform"),
    ((30, 4, 25), 2000, (4, 70), 200, "smallint"),
]


@pytest.mark.parametrize("planner", ["exact", "learned"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_fused_batch_materialized(gpu_ctx, planner, force):
    print(f"planner={planner} force={force}", flush=True)
    _load_mlp(gpu_ctx)
    pairs = _pairs(3300, PAIR_SPECS)
    rs = gpu_ctx.contract_batch(pairs, [1], [0], planner=planner, force_acc=force)
    assert len(rs) == len(pairs)
    for i, ((a, b), r) in enumerate(zip(pairs, rs)):
        ref = oracle.contract_tiled(a, (1,), b, (0,), tile_size=8)
        c, v = r.to_host()
        assert r.order == 3
        assert 0.0 < r.stats["ms_done"] <= r.stats["ms_total"] + 1e-6
        assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"pair {i}", absum=ref.absum,
                           nterms=ref.nterms)
    # free out of order: the pairs share one device block
    for r in rs[::-1]:
        r.free()


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_fused_batch_sink(gpu_ctx, planner):
    """Count + checksum per pair; the hash linearises over the bounding-box
    extents of all pairs (the oracle is given the same extents)."""
    print(f"planner={planner}", flush=True)
    _load_mlp(gpu_ctx)
    pairs = _pairs(3400, PAIR_SPECS)
    ea, eb = _bbox([p[0] for p in pairs]), _bbox([p[1] for p in pairs])
    rs = gpu_ctx.contract_batch(pairs, [1], [0], planner=planner, sink=True)
    for i, ((a, b), r) in enumerate(zip(pairs, rs)):
        ref = oracle.contract_tiled(a, (1,), b, (0,), tile_size=8, sink=True, a_extents=ea, b_extents=eb)
        assert r.nnz == ref.nnz, f"pair {i}"
        assert r.coord_hash == ref.coord_hash, f"pair {i}"
        assert abs(r.val_sum - ref.val_sum) <= 1e-9 * max(1.0, abs(ref.val_sum)), f"pair {i}"
        assert abs(r.val_wsum - ref.val_wsum) <= 1e-9 * max(1.0, abs(ref.val_wsum)), f"pair {i}"
        r.free()


def test_fused_batch_many_pairs_sink(gpu_ctx):
    """2100 small pairs in one launch: the per-pair summaries read back at the
    end (4 x 8 bytes per pair) exceed the library's pinned read-back buffer and
    take its direct-copy path; every pair still matches the oracle."""
    rng = np.random.default_rng(3500)
    specs = [((5, 4, 6), int(rng.integers(1, 40)), (4, 7), int(rng.integers(1, 20)), "smallint")
             for _ in range(2100)]
    pairs = _pairs(3501, specs)
    ea, eb = _bbox([p[0] for p in pairs]), _bbox([p[1] for p in pairs])
    rs = gpu_ctx.contract_batch(pairs, [1], [0], sink=True)
    assert len(rs) == len(pairs)
    for i, ((a, b), r) in enumerate(zip(pairs, rs)):
        ref = oracle.contract_tiled(a, (1,), b, (0,), tile_size=8, sink=True, a_extents=ea, b_extents=eb)
        assert r.nnz == ref.nnz, f"pair {i}"
        assert r.coord_hash == ref.coord_hash, f"pair {i}"
        assert r.val_sum == ref.val_sum and r.val_wsum == pytest.approx(ref.val_wsum, rel=1e-12, abs=1e-12), f"pair {i}"
        r.free()


def test_fused_batch_matches_single_calls(gpu_ctx):
    """Scaled uneven pairs of the order-4 x order-3 batch shape: each fused
    pair equals its own fcc_contract call (same coordinates, values within the
    summation-order bound)."""
    wls = W.c5_pairs(n_pairs=12, scale=0.01)
    pairs = [wl.tensors() for wl in wls]
    rs = gpu_ctx.contract_batch(pairs, [1, 3], [1, 2])
    for i, ((a, b), r) in enumerate(zip(pairs, rs)):
        print(f"pair {i}: nnzA {a.nnz} nnzB {b.nnz} out {r.nnz}", flush=True)
        s = gpu_ctx.contract(a, [1, 3], b, [1, 2])
        c1, v1 = r.to_host()
        c2, v2 = s.to_host()
        ref = oracle.contract_tiled(a, (1, 3), b, (1, 2), tile_size=128)
        assert_same_result(c1, v1, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"fused pair {i}",
                           absum=ref.absum, nterms=ref.nterms)
        assert_same_result(c2, v2, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"single pair {i}",
                           absum=ref.absum, nterms=ref.nterms)
        s.free()
        r.free()


def test_fused_batch_edge_cases(gpu_ctx):
    # no pairs
    assert gpu_ctx.contract_batch([], [1], [0]) == []
    # every pair empty on one side
    rng = np.random.default_rng(3500)
    e3 = SparseTensor(np.zeros((3, 0), np.uint32), np.zeros(0))
    pairs = [(e3, rand_tensor(rng, (4, 5), 6, "uniform")), (e3, rand_tensor(rng, (4, 5), 6, "uniform"))]
    rs = gpu_ctx.contract_batch(pairs, [1], [0])
    assert [r.nnz for r in rs] == [0, 0]
    # pairs that differ in order are rejected
    a2 = rand_tensor(rng, (4, 5), 5, "uniform")
    a3 = rand_tensor(rng, (4, 5, 3), 5, "uniform")
    b2 = rand_tensor(rng, (5, 3), 5, "uniform")
    with pytest.raises(FastccError):
        gpu_ctx.contract_batch([(a2, b2), (a3, b2)], [1], [0])
