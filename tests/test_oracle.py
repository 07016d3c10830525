"""CPU tests pinning the oracle (oracle/fcc_oracle.c).

The reference ships no tests, fixtures or golden vectors, and cannot be
built here (its five dependency submodules are empty), so the oracle is pinned
by what the mathematics fixes:
  - brute force == dense numpy einsum on densified tiny tensors (the README's
    stated semantics result(i,j,k) = a(i,c0,j,c1) * b(k,c0,c1), README.md:8-13),
    support == boolean einsum (structural, exact zeros kept);
  - the tiled restatement == brute force, over tile sizes, thread counts and
    mode lists (bit-exact with small-integer values);
  - all-ones inputs give integer match counts;
  - the reference's hand-checkable case of its decode defect (index_tensors.hpp:1042);
  - golden vectors committed under tests/golden (written by the oracle,
    cross-checked here against einsum) so a later change to the oracle shows.
"""
import itertools
import json
import os
import string

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, densify, rand_tensor, sort_result

from oracle import oracle
from fastcc_amd import workloads as W
from fastcc_amd.workloads import SparseTensor


def einsum_contract(a, am, b, bm, Ea, Eb, boolean=False, ab=(), bb=()):
    """Dense reference with numpy: output modes batch pairs (list order), then
    A's other externals, then B's (einsum keeps a letter shared by both
    operands and named in the output as a batch index)."""
    letters = iter(string.ascii_letters)
    la = [None] * a.order
    lb = [None] * b.order
    batch = []
    for x, y in zip(ab, bb):
        c = next(letters)
        la[x] = c
        lb[y] = c
        batch.append(c)
    for x, y in zip(am, bm):
        c = next(letters)
        la[x] = c
        lb[y] = c
    for i in range(a.order):
        if la[i] is None:
            la[i] = next(letters)
    for i in range(b.order):
        if lb[i] is None:
            lb[i] = next(letters)
    out = batch + [la[i] for i in range(a.order) if i not in am and i not in ab] + \
        [lb[i] for i in range(b.order) if i not in bm and i not in bb]
    # paired extents share the larger radix
    Ea, Eb = list(Ea), list(Eb)
    for x, y in list(zip(am, bm)) + list(zip(ab, bb)):
        m = max(Ea[x], Eb[y])
        Ea[x] = Eb[y] = m
    da, db = densify(a, tuple(Ea)), densify(b, tuple(Eb))
    if boolean:
        da, db = (da != 0) | _struct(a, Ea), (db != 0) | _struct(b, Eb)
        da, db = da.astype(np.int64), db.astype(np.int64)
    expr = "".join(la) + "," + "".join(lb) + "->" + "".join(out)
    return np.einsum(expr, da, db)


def _struct(t, dims):
    s = np.zeros(dims, dtype=bool)
    if t.nnz:
        s[tuple(t.coords.astype(np.int64))] = True
    return s


def dense_to_coo(count, val):
    if count.ndim == 0:
        n = int(count != 0)
        return np.zeros((0, n), dtype=np.uint32), np.array([float(val)] * n)
    nz = np.nonzero(count)
    coords = np.array(nz, dtype=np.uint32).reshape(count.ndim, -1)
    return coords, val[nz]


CASES = [
    # (dims A, nnz A, a_modes, dims B, nnz B, b_modes)
    ((5, 4, 6, 3), 40, (1, 3), (4, 4, 3), 25, (1, 2)),   # README shape: a(i,c0,j,c1) b(k,c0,c1)
    ((6, 5, 4), 50, (0,), (6, 7), 20, (0,)),
    ((3, 4, 5), 30, (2, 0), (5, 6, 3), 40, (0, 2)),       # non-leading, permuted pairs
    ((4, 5), 12, (), (3, 2), 5, ()),                      # outer product (no contraction)
    ((4, 5), 12, (0, 1), (4, 5), 9, (0, 1)),              # full contraction: order-0 output
    ((7, 3, 2, 4), 60, (1,), (3, 5), 10, (0,)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("values", ["uniform", "smallint", "zeros"])
def test_bruteforce_matches_einsum(case, values):
    rng = np.random.default_rng(100 + case)
    da, na, am, db, nb, bm = CASES[case]
    a = rand_tensor(rng, da, na, values)
    b = rand_tensor(rng, db, nb, values)
    res = oracle.contract_bruteforce(a, am, b, bm)
    Ea, Eb = a.extents(), b.extents()
    cnt = einsum_contract(a, am, b, bm, Ea, Eb, boolean=True)
    val = einsum_contract(a, am, b, bm, Ea, Eb)
    rc, rv = dense_to_coo(cnt, val)
    assert res.nnz == rc.shape[1]
    assert_same_result(res.coords, res.vals, rc, rv, rel=1e-13, abs_floor=1e-13, what="bruteforce vs einsum")


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("tile,threads", [(1, 1), (2, 3), (100, 2), (128, 8), (4096, 4)])
def test_tiled_matches_bruteforce(case, tile, threads):
    rng = np.random.default_rng(200 + case)
    da, na, am, db, nb, bm = CASES[case]
    a = rand_tensor(rng, da, na, "smallint")
    b = rand_tensor(rng, db, nb, "smallint")
    ref = oracle.contract_bruteforce(a, am, b, bm)
    for hp in (True, False):
        got = oracle.contract_tiled(a, am, b, bm, tile_size=tile, threads=threads, high_precision=hp)
        # small-integer values: every sum is exact, so equality is exact
        assert_same_result(got.coords, got.vals, ref.coords, ref.vals, rel=0, abs_floor=0, what="tiled")
        assert got.products == ref.products


BATCH_CASES = [
    # (dims A, nnz A, a_modes, a_batch, dims B, nnz B, b_modes, b_batch)
    ((4, 5, 6), 60, (2,), (0,), (4, 6, 3), 40, (1,), (0,)),          # bij,bjk->bik
    ((3, 5, 4, 6), 90, (1, 3), (2,), (6, 4, 5, 2), 70, (3, 0), (1,)),  # batch mode in the middle
    ((3, 4, 5), 40, (), (0, 2), (5, 3, 6), 30, (), (0, 1)),           # batch only, two batch pairs
    ((2, 3, 4), 20, (1,), (2, 0), (3, 2, 4), 22, (0,), (2, 1)),       # every A mode keyed
]


@pytest.mark.parametrize("case", range(len(BATCH_CASES)))
@pytest.mark.parametrize("values", ["uniform", "zeros"])
def test_batch_bruteforce_matches_einsum(case, values):
    """Batch pairs (contract.hpp:285-367): coordinates must agree like a
    contraction pair and lead the output (index_tensors.hpp:318-353)."""
    rng = np.random.default_rng(700 + case)
    da, na, am, ab, db, nb, bm, bb = BATCH_CASES[case]
    a = rand_tensor(rng, da, na, values)
    b = rand_tensor(rng, db, nb, values)
    res = oracle.contract_bruteforce(a, am, b, bm, a_batch=ab, b_batch=bb)
    Ea, Eb = a.extents(), b.extents()
    cnt = einsum_contract(a, am, b, bm, Ea, Eb, boolean=True, ab=ab, bb=bb)
    val = einsum_contract(a, am, b, bm, Ea, Eb, ab=ab, bb=bb)
    rc, rv = dense_to_coo(cnt, val)
    assert res.order == len(ab) + a.order + b.order - 2 * (len(am) + len(ab))
    assert res.nnz == rc.shape[1]
    assert_same_result(res.coords, res.vals, rc, rv, rel=1e-13, abs_floor=1e-13, what="batch bruteforce vs einsum")


@pytest.mark.parametrize("case", range(len(BATCH_CASES)))
@pytest.mark.parametrize("tile,threads", [(1, 1), (4, 3), (128, 8)])
def test_batch_tiled_matches_bruteforce(case, tile, threads):
    rng = np.random.default_rng(800 + case)
    da, na, am, ab, db, nb, bm, bb = BATCH_CASES[case]
    a = rand_tensor(rng, da, na, "smallint")
    b = rand_tensor(rng, db, nb, "smallint")
    ref = oracle.contract_bruteforce(a, am, b, bm, a_batch=ab, b_batch=bb)
    got = oracle.contract_tiled(a, am, b, bm, tile_size=tile, threads=threads, a_batch=ab, b_batch=bb)
    assert_same_result(got.coords, got.vals, ref.coords, ref.vals, rel=0, abs_floor=0, what="batch tiled")
    assert got.products == ref.products


def test_batch_equals_per_slice_contractions():
    """A batch pair is the same as contracting each batch slice on its own and
    prefixing the slice index: checked against the plain (batch-free) oracle."""
    rng = np.random.default_rng(901)
    a = rand_tensor(rng, (5, 7, 6), 120, "uniform")
    b = rand_tensor(rng, (6, 5, 4), 80, "uniform")
    got = oracle.contract_bruteforce(a, (2,), b, (0,), a_batch=(0,), b_batch=(1,))
    cs, vs = [], []
    for z in range(5):
        ka, kb = a.coords[0] == z, b.coords[1] == z
        sa = SparseTensor(np.ascontiguousarray(a.coords[1:, ka]), a.vals[ka])
        sb = SparseTensor(np.ascontiguousarray(b.coords[[0, 2]][:, kb]), b.vals[kb])
        if sa.nnz == 0 or sb.nnz == 0:
            continue
        # pin the slice extents to the full ones so coordinates line up
        r = oracle.contract_bruteforce(sa, (1,), sb, (0,), a_extents=(7, 6), b_extents=(6, 4))
        cs.append(np.vstack([np.full((1, r.nnz), z, dtype=np.uint32), r.coords]))
        vs.append(r.vals)
    rc, rv = np.hstack(cs), np.concatenate(vs)
    assert_same_result(got.coords, got.vals, rc, rv, rel=1e-15, abs_floor=0, what="batch vs slices")


def test_explicit_extents_checked():
    rng = np.random.default_rng(902)
    a = rand_tensor(rng, (5, 6), 20, "uniform")
    b = rand_tensor(rng, (6, 3), 10, "uniform")
    r1 = oracle.contract_tiled(a, (1,), b, (0,), sink=True)
    r2 = oracle.contract_tiled(a, (1,), b, (0,), sink=True, a_extents=(9, 6), b_extents=(6, 3))
    assert r1.nnz == r2.nnz and r1.val_sum == r2.val_sum
    with pytest.raises(RuntimeError):
        oracle.contract_tiled(a, (1,), b, (0,), a_extents=(2, 6))


def test_all_ones_counts_matches():
    rng = np.random.default_rng(7)
    a = rand_tensor(rng, (9, 6, 8), 200, "ones")
    b = rand_tensor(rng, (6, 8, 5), 120, "ones")
    got = oracle.contract_tiled(a, (1, 2), b, (0, 1), tile_size=4, threads=2)
    cnt = einsum_contract(a, (1, 2), b, (0, 1), a.extents(), b.extents(), boolean=True)
    rc, rv = dense_to_coo(cnt, cnt.astype(np.float64))
    assert_same_result(got.coords, got.vals, rc, rv, rel=0, abs_floor=0, what="ones")
    assert np.all(got.vals == np.round(got.vals)) and got.vals.min() >= 1


def test_reference_decode_defect_example():
    """T=2, A={(1,0):1}, B={(0,0):1}, contract A mode 1 with B mode 1: the
    correct result is C(1,0)=1 (the reference's natural-log shift would emit
    C(0,0), index_tensors.hpp:1042)."""
    from fastcc_amd.workloads import SparseTensor
    a = SparseTensor(np.array([[1], [0]], dtype=np.uint32), np.array([1.0]))
    b = SparseTensor(np.array([[0], [0]], dtype=np.uint32), np.array([1.0]))
    got = oracle.contract_tiled(a, (1,), b, (1,), tile_size=2, threads=1)
    assert got.nnz == 1
    assert got.coords[:, 0].tolist() == [1, 0] and got.vals[0] == 1.0


def test_structural_zero_kept():
    """Exact-zero sums stay in the output (index_tensors.hpp:1028-1037)."""
    from fastcc_amd.workloads import SparseTensor
    a = SparseTensor(np.array([[0, 0], [0, 1]], dtype=np.uint32), np.array([1.0, 1.0]))
    b = SparseTensor(np.array([[0, 1], [0, 0]], dtype=np.uint32), np.array([1.0, -1.0]))
    for res in (oracle.contract_bruteforce(a, (1,), b, (0,)), oracle.contract_tiled(a, (1,), b, (0,), tile_size=2)):
        assert res.nnz == 1 and res.vals[0] == 0.0


def test_sink_summaries_match_materialised():
    rng = np.random.default_rng(11)
    a = rand_tensor(rng, (20, 6, 15), 300)
    b = rand_tensor(rng, (6, 25), 80)
    full = oracle.contract_tiled(a, (1,), b, (0,), tile_size=8, threads=3)
    sk = oracle.contract_tiled(a, (1,), b, (0,), tile_size=8, threads=3, sink=True)
    assert sk.nnz == full.nnz and sk.coord_hash == full.coord_hash
    assert sk.val_sum == pytest.approx(full.val_sum, rel=1e-12, abs=1e-12)
    assert sk.val_wsum == pytest.approx(full.val_wsum, rel=1e-12, abs=1e-12)
    # summary definition: sum of mix64(linear output index)
    Ea, Eb = a.extents(), b.extents()
    lin = (full.coords[0].astype(np.uint64) * np.uint64(Ea[2]) + full.coords[1].astype(np.uint64))
    lin = lin * np.uint64(Eb[1]) + full.coords[2].astype(np.uint64)
    h = 0
    for x in lin.tolist():
        h = (h + oracle.mix64(x)) & 0xFFFFFFFFFFFFFFFF
    assert h == full.coord_hash


def test_sampled_is_subset():
    rng = np.random.default_rng(12)
    a = rand_tensor(rng, (64, 5, 9), 700)
    b = rand_tensor(rng, (5, 9, 30), 600)
    full = oracle.contract_tiled(a, (1, 2), b, (0, 1), tile_size=4, threads=2)
    samp = oracle.contract_tiled(a, (1, 2), b, (0, 1), tile_size=4, threads=2, sample_stride=3)
    keep = (full.coords[0] // 4) % 3 == 0
    assert_same_result(samp.coords, samp.vals, full.coords[:, keep], full.vals[keep], rel=0, abs_floor=0)


def test_generator_deterministic_unique_and_skewed():
    spec = W.TensorSpec((50, 40, 30), 5000, zipf={0: 1.2}, value_dist=W.LOGNORMAL, seed=123)
    t1, t2 = W.generate(spec), W.generate(spec)
    assert np.array_equal(t1.coords, t2.coords) and np.array_equal(t1.vals, t2.vals)
    lin = (t1.coords[0].astype(np.int64) * 40 + t1.coords[1]) * 30 + t1.coords[2]
    assert np.unique(lin).size == 5000
    h = np.bincount(t1.coords[0], minlength=50)
    assert h[0] > h[10] > h[40]          # Zipf head
    assert np.all(t1.vals > 0)            # lognormal
    u = W.generate(W.TensorSpec((1000,), 900, value_dist=W.UNIFORM_PM1, seed=5))
    assert u.vals.min() >= -1 and u.vals.max() < 1


# ------------------------------------------------------------ golden vectors

GOLDEN = os.path.join(ROOT, "tests", "golden", "oracle_small.json")


def test_golden_vectors():
    """Committed vectors (inputs + outputs) still come out of the oracle, and
    agree with numpy einsum."""
    with open(GOLDEN) as f:
        cases = json.load(f)
    from fastcc_amd.workloads import SparseTensor
    for c in cases:
        a = SparseTensor(np.array(c["a_coords"], dtype=np.uint32), np.array(c["a_vals"]))
        b = SparseTensor(np.array(c["b_coords"], dtype=np.uint32), np.array(c["b_vals"]))
        got = oracle.contract_tiled(a, c["a_modes"], b, c["b_modes"], tile_size=c["tile"], threads=2)
        ec = np.array(c["c_coords"], dtype=np.uint32).reshape(len(c["c_coords"]), -1)
        assert_same_result(got.coords, got.vals, ec, np.array(c["c_vals"]), rel=0, abs_floor=0, what=c["name"])
        cnt = einsum_contract(a, c["a_modes"], b, c["b_modes"], a.extents(), b.extents(), boolean=True)
        val = einsum_contract(a, c["a_modes"], b, c["b_modes"], a.extents(), b.extents())
        rc, rv = dense_to_coo(cnt, val)
        assert_same_result(ec, np.array(c["c_vals"]), rc, rv, rel=0, abs_floor=0, what=c["name"] + " einsum")


def test_absum_and_nterms():
    """Per-output sum of |terms| and term count (the summation-order tolerance)."""
    rng = np.random.default_rng(21)
    a = rand_tensor(rng, (8, 5, 7), 150)
    b = rand_tensor(rng, (5, 7, 9), 140)
    for res in (oracle.contract_tiled(a, (1, 2), b, (0, 1), tile_size=4, threads=2),
                oracle.contract_bruteforce(a, (1, 2), b, (0, 1))):
        Ea, Eb = a.extents(), b.extents()
        cnt = einsum_contract(a, (1, 2), b, (0, 1), Ea, Eb, boolean=True)
        aa = type(a)(a.coords, np.abs(a.vals))
        bb = type(b)(b.coords, np.abs(b.vals))
        ab = einsum_contract(aa, (1, 2), bb, (0, 1), Ea, Eb)
        rc, rab = dense_to_coo(cnt, ab)
        _, rn = dense_to_coo(cnt, cnt.astype(np.float64))
        o = np.lexsort(res.coords[::-1])
        assert np.array_equal(res.coords[:, o], rc)
        np.testing.assert_allclose(res.absum[o], rab, rtol=1e-12)
        assert np.array_equal(res.nterms[o], rn.astype(np.uint32))
