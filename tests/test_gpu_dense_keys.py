"""GPU: contraction keys taken by the dense FP64 GEMM (fcc_options.dense_keys).

A key whose left and right fibers are both dense leaves the tile traversal;
all such keys are contracted together by one matrix-core GEMM and every unit
starts its accumulator from that block, presence kept exact by key bitsets.
The outputs must not change: coordinate sets bit-exact against the oracle,
values within the fp64 tolerance (exactly, for small-integer inputs). Each
test checks that the path was taken (stats dense_keys > 0)."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_GHASH, FCC_ACC_LHASH
from fastcc_amd import workloads as W
from fastcc_amd.workloads import SparseTensor

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14
ALL_KEYS = 1e-9  # every key with entries on both sides

# (A dims, nnz, A modes, B dims, nnz, B modes): spans off the GEMM tile
# (128), key counts off its k step (16), one and several k steps and tiles
CASES = [
    ((5, 4, 6, 3), 40, (1, 3), (4, 4, 3), 25, (1, 2)),
    ((30, 7, 11), 1200, (1,), (7, 40), 200, (0,)),
    ((300, 37), 6000, (1,), (37, 200), 5000, (0,)),
    ((150, 500), 20000, (1,), (500, 170), 20000, (0,)),
    ((9, 13, 20), 1500, (1, 0), (13, 9, 150), 9000, (0, 1)),
]


def _check(r, ref, exact, what):
    c, v = r.to_host()
    assert_same_result(c, v, ref.coords, ref.vals, rel=0 if exact else REL, abs_floor=0 if exact else ABS,
                       what=what, absum=None if exact else ref.absum, nterms=None if exact else ref.nterms)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("values", ["uniform", "smallint", "zeros"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_all_keys_dense(gpu_ctx, case, values, force):
    print(f"case={case} values={values} force={force}", flush=True)
    rng = np.random.default_rng(4200 + case)
    da, na, am, db, nb, bm = CASES[case]
    a = rand_tensor(rng, da, na, values)
    b = rand_tensor(rng, db, nb, values)
    ref = oracle.contract_bruteforce(a, am, b, bm)
    r = gpu_ctx.contract(a, list(am), b, list(bm), force_acc=force, dense_keys=ALL_KEYS)
    st = r.stats
    assert st["dense_keys"] > 0
    assert st["dense_key_products"] == st["products"] == ref.products
    # only dense units can be counted from the block
    assert st["units_block_counted"] <= st["units_dense"]
    if force in (FCC_ACC_LHASH, FCC_ACC_GHASH):
        assert st["units_block_counted"] == 0
    _check(r, ref, values != "uniform", f"case {case} {values} force {force}")
    r.free()


@pytest.mark.parametrize("tau", [0.002, 0.02, 0.1])
@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_mixed_dense_and_sparse(gpu_ctx, tau, planner):
    """Zipf keys: some keys dense, the rest through the tiles."""
    print(f"tau={tau} planner={planner}", flush=True)
    gpu_ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    wl = W.c2(0.05)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=128, high_precision=True)
    r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), planner=planner, dense_keys=tau)
    st = r.stats
    assert 0 < st["dense_key_products"] < st["products"]
    _check(r, ref, False, f"c2 tau {tau} {planner}")
    r.free()


@pytest.mark.parametrize("name,mk", [("c2", lambda: W.c2(0.25)), ("c4", lambda: W.c4(0.05))])
@pytest.mark.parametrize("sink", [False, True])
def test_default_threshold(gpu_ctx, name, mk, sink):
    """The default rule picks dense keys on the Zipf workloads; the summary
    (count, coordinate hash, value sums) and the outputs match the oracle."""
    wl = mk()
    print(f"workload={wl.name} sink={sink}", flush=True)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=128, high_precision=True, sink=sink)
    r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), sink=sink, summarize=True)
    assert r.stats["dense_keys"] > 0
    assert r.stats["units_block_counted"] <= r.stats["units_dense"]
    assert r.nnz == ref.nnz and r.coord_hash == ref.coord_hash
    assert r.val_sum == pytest.approx(ref.val_sum, rel=1e-10, abs=1e-10)
    assert r.val_wsum == pytest.approx(ref.val_wsum, rel=1e-10, abs=1e-10)
    if not sink:
        _check(r, ref, False, wl.name)
    off = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), sink=True, dense_keys=-1)
    assert off.stats["dense_keys"] == 0
    assert off.nnz == r.nnz and off.coord_hash == r.coord_hash
    off.free()
    r.free()


@pytest.mark.parametrize("values", ["uniform", "smallint"])
@pytest.mark.parametrize("shape", [((40, 30), (30, 50)), ((300, 37), (37, 200)), ((9, 13, 20), (13, 9, 150))])
def test_block_covered_units_counted_from_block(gpu_ctx, values, shape):
    """Fully stored operands: every output cell is reached by a dense key, so
    the exact planner counts every dense unit from the block's presence map
    instead of traversing it; sizes and values still match the oracle."""
    print(f"shape={shape} values={values}", flush=True)
    rng = np.random.default_rng(4500 + len(shape[0]))
    da, db = shape
    nA, nB = int(np.prod(da)), int(np.prod(db))
    a = rand_tensor(rng, da, nA, values)  # every coordinate stored
    b = rand_tensor(rng, db, nB, values)
    am = (1,) if len(da) == 2 else (1, 0)
    bm = (0,) if len(db) == 2 else (0, 1)
    ref = oracle.contract_bruteforce(a, am, b, bm)
    r = gpu_ctx.contract(a, list(am), b, list(bm), dense_keys=ALL_KEYS, force_acc=FCC_ACC_DENSE)
    st = r.stats
    assert st["dense_keys"] > 0 and st["units_dense"] > 0
    assert st["units_block_counted"] == st["units_dense"]
    assert r.nnz == ref.nnz
    _check(r, ref, values != "uniform", f"covered {shape} {values}")
    r.free()


# Self-contractions (one tensor, the same modes on both sides): the GEMM's
# operands are one panel, so the block is symmetric; it computes the tiles on
# and above the diagonal and mirrors them. Spans of one tile, a partial band
# of tiles (3, 5), two bands (9 tiles: 8 + 1); 300 keys over 5 tiles
# (15 computed) give a split k.
SELF_CASES = [
    ((40, 90), 3000, (0,)),
    ((300, 600), 20000, (0,)),
    ((20, 15, 30), 6000, (1,)),
    ((60, 1100), 30000, (0,)),
    ((6, 5, 40, 9), 5000, (0, 1)),
]


@pytest.mark.parametrize("case", range(len(SELF_CASES)))
@pytest.mark.parametrize("values", ["uniform", "smallint"])
def test_self_contraction_symmetric_block(gpu_ctx, case, values):
    print(f"case={case} values={values}", flush=True)
    rng = np.random.default_rng(4600 + case)
    dims, nnz, modes = SELF_CASES[case]
    a = rand_tensor(rng, dims, nnz, values)
    ref = oracle.contract_bruteforce(a, modes, a, modes)
    r = gpu_ctx.contract(a, list(modes), a, list(modes), dense_keys=ALL_KEYS)
    st = r.stats
    assert st["dense_keys"] > 0
    assert st["dense_key_products"] == st["products"] == ref.products
    span = int(np.prod([d for m, d in enumerate(dims) if m not in modes]))
    nbm, kp = -(-span // 128), -(-st["dense_keys"] // 16) * 16
    assert st["dense_flops"] == 2 * (nbm * (nbm + 1) // 2) * 128 * 128 * kp  # the upper triangle of tiles
    _check(r, ref, values != "uniform", f"self case {case} {values}")
    r.free()


@pytest.mark.parametrize("use_graph", [False, True])
def test_self_plan_new_values(gpu_ctx, use_graph):
    """A self-contraction plan re-runs the half GEMM on new values (two tiles
    of rows, so one mirrored tile)."""
    rng = np.random.default_rng(4700)
    a = rand_tensor(rng, (50, 200), 4000, "uniform")
    plan = gpu_ctx.plan(a, [0], a, [0], dense_keys=ALL_KEYS)
    for it in range(3):
        a2 = SparseTensor(a.coords, rng.uniform(-1, 1, a.nnz))
        ref = oracle.contract_bruteforce(a2, (0,), a2, (0,))
        r = plan.execute(a2.vals, graph=use_graph)
        assert r.stats["dense_keys"] > 0
        _check(r, ref, False, f"self plan it {it} graph {use_graph}")
    plan.close()


def test_batch_modes(gpu_ctx):
    """Batch pairs lead rows and keys: a key is dense within its batch."""
    rng = np.random.default_rng(4300)
    a = rand_tensor(rng, (3, 20, 6, 9), 2500, "smallint")   # a(b, i, k, j)
    b = rand_tensor(rng, (6, 3, 30), 400, "smallint")       # b(k, b, l)
    ref = oracle.contract_bruteforce(a, (2,), b, (0,), a_batch=(0,), b_batch=(1,))
    r = gpu_ctx.contract(a, [2], b, [0], a_batch=[0], b_batch=[1], dense_keys=ALL_KEYS)
    assert r.stats["dense_keys"] > 0
    _check(r, ref, True, "batch")
    r.free()


@pytest.mark.parametrize("use_graph", [False, True])
def test_plan_new_values(gpu_ctx, use_graph):
    """A plan re-scatters the dense panels and re-runs the GEMM on new values."""
    rng = np.random.default_rng(4400)
    a = rand_tensor(rng, (90, 40), 2500, "uniform")
    b = rand_tensor(rng, (40, 70), 2000, "uniform")
    plan = gpu_ctx.plan(a, [1], b, [0], dense_keys=ALL_KEYS)
    for it in range(3):
        a2 = SparseTensor(a.coords, rng.uniform(-1, 1, a.nnz))
        b2 = SparseTensor(b.coords, rng.uniform(-1, 1, b.nnz))
        ref = oracle.contract_bruteforce(a2, (1,), b2, (0,))
        r = plan.execute(a2.vals, b2.vals, graph=use_graph)
        _check(r, ref, False, f"plan it {it} graph {use_graph}")
    plan.close()


def test_sampled_left_rows(gpu_ctx):
    """Sampled left tiles (the large-shape baseline mode) with dense keys:
    the keys are counted from the sampled rows only."""
    wl = W.c2(0.05)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=8, sample_stride=3)
    r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), tile_hint=8, left_tile_stride=3,
                         dense_keys=0.005)
    assert r.stats["dense_keys"] > 0
    _check(r, ref, False, "sampled dense keys")
    r.free()
