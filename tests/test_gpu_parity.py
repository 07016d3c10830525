"""GPU parity: the HIP contraction against the CPU oracle on the same seeded
inputs. Coordinates must match bit-exactly as sorted sets; values within
max(1e-12 * |ref|, 1e-14) (small-integer inputs: exactly).

Parity is against the contraction's stated semantics (README.md:8-13), which
the oracle restates from contract.hpp:824-912 with the reference's decode
defects corrected (see oracle/fcc_oracle.c)."""
import numpy as np
import pytest

from conftest import assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_GHASH, FCC_ACC_LHASH, FastccError
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14

SMALL = [
    ((5, 4, 6, 3), 40, (1, 3), (4, 4, 3), 25, (1, 2)),
    ((6, 5, 4), 50, (0,), (6, 7), 20, (0,)),
    ((3, 4, 5), 30, (2, 0), (5, 6, 3), 40, (0, 2)),
    ((4, 5), 12, (), (3, 2), 5, ()),
    ((4, 5), 12, (0, 1), (4, 5), 9, (0, 1)),
    ((7, 3, 2, 4), 60, (1,), (3, 5), 10, (0,)),
    ((40, 9, 30), 2000, (1,), (9, 50), 300, (0,)),
    ((300, 7, 8), 6000, (1, 2), (7, 8, 200), 3000, (0, 1)),
]


def gpu_contract(ctx, a, am, b, bm, **kw):
    r = ctx.contract(a, list(am), b, list(bm), **kw)
    c, v = r.to_host()
    st = r.stats
    r.free()
    return c, v, st


@pytest.mark.parametrize("case", range(len(SMALL)))
@pytest.mark.parametrize("values", ["uniform", "smallint", "zeros"])
@pytest.mark.parametrize("force", [FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_LHASH, FCC_ACC_GHASH])
def test_small_cases(gpu_ctx, case, values, force):
    print(f"case={case} values={values} force={force}", flush=True)
    rng = np.random.default_rng(1000 + case)
    da, na, am, db, nb, bm = SMALL[case]
    a = rand_tensor(rng, da, na, values)
    b = rand_tensor(rng, db, nb, values)
    ref = oracle.contract_bruteforce(a, am, b, bm)
    c, v, st = gpu_contract(gpu_ctx, a, am, b, bm, force_acc=force)
    exact = values != "uniform"
    assert_same_result(c, v, ref.coords, ref.vals, rel=0 if exact else REL, abs_floor=0 if exact else ABS,
                       what=f"case {case} {values} force {force}", absum=None if exact else ref.absum,
                       nterms=None if exact else ref.nterms)
    assert st["products"] == ref.products


@pytest.mark.parametrize("tile", [0, 1, 2, 16, 100, 128, 2048])
def test_tile_hints(gpu_ctx, tile):
    print(f"tile={tile}", flush=True)
    rng = np.random.default_rng(7)
    a = rand_tensor(rng, (60, 12, 40), 5000, "smallint")
    b = rand_tensor(rng, (12, 70), 400, "smallint")
    ref = oracle.contract_bruteforce(a, (1,), b, (0,))
    c, v, st = gpu_contract(gpu_ctx, a, (1,), b, (0,), tile_hint=tile)
    assert_same_result(c, v, ref.coords, ref.vals, rel=0, abs_floor=0, what=f"tile {tile}")


SCALED = [
    ("c1", lambda: W.c1(0.5)),
    ("c2", lambda: W.c2(0.25)),
    ("c3", lambda: W.c3(0.02)),
    ("c4", lambda: W.c4(0.05)),
]


@pytest.mark.parametrize("name,mk", SCALED)
@pytest.mark.parametrize("planner_sink", [("exact", False), ("exact", True)])
def test_scaled_workloads(gpu_ctx, name, mk, planner_sink):
    planner, sink = planner_sink
    wl = mk()
    print(f"workload={wl.name} planner={planner} sink={sink}", flush=True)
    a, b = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, b, wl.b_modes, tile_size=128, high_precision=True, sink=sink)
    r = gpu_ctx.contract(a, list(wl.a_modes), b if not wl.self_contraction else a, list(wl.b_modes), sink=sink,
                         summarize=True)
    assert r.nnz == ref.nnz
    assert r.coord_hash == ref.coord_hash
    assert r.val_sum == pytest.approx(ref.val_sum, rel=1e-10, abs=1e-10)
    assert r.val_wsum == pytest.approx(ref.val_wsum, rel=1e-10, abs=1e-10)
    if not sink:
        c, v = r.to_host()
        assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=wl.name, absum=ref.absum,
                           nterms=ref.nterms)
    assert r.stats["products"] == ref.products
    r.free()


def test_sampled_left_tiles(gpu_ctx):
    wl = W.c3(0.05)
    a, _ = wl.tensors()
    ref = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=64, sample_stride=16)
    r = gpu_ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), tile_hint=64, left_tile_stride=16)
    c, v = r.to_host()
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what="sampled", absum=ref.absum,
                       nterms=ref.nterms)
    r.free()


def test_cross_kind_invariance(gpu_ctx):
    """The same tiles forced through each accumulator kind give identical
    coordinate sets and (small-integer inputs) identical values."""
    wl = W.c2(0.03)
    a, _ = wl.tensors()
    a.vals = np.round(a.vals * 4) + 1.0
    outs = []
    for force in (FCC_ACC_DENSE, FCC_ACC_LHASH, FCC_ACC_GHASH, FCC_ACC_AUTO):
        c, v, st = gpu_contract(gpu_ctx, a, wl.a_modes, a, wl.b_modes, force_acc=force)
        outs.append((c, v))
    for c, v in outs[1:]:
        assert_same_result(c, v, outs[0][0], outs[0][1], rel=0, abs_floor=0, what="kinds")


def test_empty_and_disjoint(gpu_ctx):
    from fastcc_amd.workloads import SparseTensor
    e = SparseTensor(np.zeros((3, 0), dtype=np.uint32), np.zeros(0))
    rng = np.random.default_rng(3)
    b = rand_tensor(rng, (4, 5, 6), 30)
    r = gpu_ctx.contract(e, [1], b, [0])
    assert r.nnz == 0 and r.order == 4
    # keys that never match; a larger contraction first leaves the library's
    # cached device blocks holding data, so a buffer the empty left side never
    # writes would be read as garbage (regression: its one-entry offset arrays)
    big = rand_tensor(rng, (300, 40, 50), 60000)
    gpu_ctx.contract(big, [1], big, [1]).free()
    a = SparseTensor(np.array([[0, 1], [7, 7]], dtype=np.uint32), np.array([1.0, 2.0]))
    bb = SparseTensor(np.array([[0, 1], [3, 2]], dtype=np.uint32), np.array([1.0, 2.0]))
    for _ in range(3):
        r = gpu_ctx.contract(a, [1], bb, [0])
        assert r.nnz == 0
        r.free()


def test_invalid_arguments(gpu_ctx):
    from fastcc_amd import Context
    rng = np.random.default_rng(4)
    a = rand_tensor(rng, (4, 5), 10)
    with pytest.raises(FastccError) as ei:
        gpu_ctx.contract(a, [2], a, [0])
    assert ei.value.code == -1
    with pytest.raises(FastccError):
        gpu_ctx.contract(a, [0, 0], a, [0, 1])
    fresh = Context(0)
    with pytest.raises(FastccError):
        fresh.contract(a, [0], a, [0], planner="learned")  # no MLP parameters yet
    fresh.close()


def test_readme_example_binary(tmp_path):
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "fastcc-benchmark_amd", "build", "fastcc_readme")
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    nnz = int(out.stdout.split()[1])
    s = float(out.stdout.split()[3])
    # recompute with the oracle from the files the program wrote (read like read.hpp)
    from fastcc_amd.workloads import SparseTensor

    def read(p):
        rows = [l.split() for l in open(p) if l[:1].isdigit()]
        c = np.array([[int(x) for x in r[:-1]] for r in rows], dtype=np.uint32).T.copy()
        return SparseTensor(c, np.array([float(r[-1]) for r in rows]))
    a, b = read(tmp_path / "a.tns"), read(tmp_path / "b.tns")
    ref = oracle.contract_bruteforce(a, (1, 3), b, (1, 2))
    assert nnz == ref.nnz
    assert s == pytest.approx(float(np.sum(ref.vals)), rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("planner", ["exact", "learned"])
@pytest.mark.parametrize("sink", [False, True])
def test_direct_tiles(gpu_ctx, planner, sink):
    """Tiles whose entries sit on distinct rows write every product straight
    out (no accumulator). A sparse left operand (most rows one entry) mixes such
    tiles with accumulated ones; the result must not change."""
    import os
    from conftest import ROOT
    print(f"planner={planner} sink={sink}", flush=True)
    gpu_ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    rng = np.random.default_rng(5150)
    a = rand_tensor(rng, (400, 300, 20), 9000, "uniform")   # rows (0, 1): 120k, ~0.07 entries per row
    b = rand_tensor(rng, (20, 500), 4000, "uniform")
    for th in (1, 4, 64):
        r = gpu_ctx.contract(a, [2], b, [0], planner=planner, sink=sink, tile_hint=th)
        assert r.stats["units_direct"] > 0, th
        if sink:
            ref = oracle.contract_tiled(a, (2,), b, (0,), tile_size=64, sink=True)
            assert r.nnz == ref.nnz and r.coord_hash == ref.coord_hash
            assert abs(r.val_sum - ref.val_sum) <= 1e-9 * max(1.0, abs(ref.val_sum))
        else:
            ref = oracle.contract_bruteforce(a, (2,), b, (0,))
            c, v = r.to_host()
            assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"direct tile_hint {th}",
                               absum=ref.absum, nterms=ref.nterms)
        r.free()
