"""Parity at the measured shapes' full sizes.

* The order-4 x order-3 uniform shape (C1, 9.2M outputs) and the order-4 Zipf
  self-contraction (C2, 6.07M dense outputs) whole, against the oracle, for
  both planners: coordinate sets bit-exact, values within max(1e-12 |ref|,
  1e-14) (outputs whose sums cancel may meet the fp64 summation-order bound
  (n + 1) 2^-53 sum|terms| instead; how many did is recorded).
* The hash-heavy order-4 shape (C3, 3.3e10 outputs) and the order-5 shape
  (C4, 3.17e8 outputs) at full shape with sampled left tiles against the
  oracle's same sample, both planners; and whole, on the GPU alone, with
  mirroring on and off (every output computed directly): the same output
  count and coordinate hash, value summaries within 1e-9.
* The batched shape (C5 pair shapes at 10% of their nonzeros, 48 pairs) in
  one fused launch with the learned planner: every pair's count and
  checksums against the oracle.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, record_parity

from oracle import oracle
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14
MLP = os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")


@pytest.fixture(scope="module")
def ctx(gpu_ctx):
    gpu_ctx.set_mlp_params(np.load(MLP))
    return gpu_ctx


_cache = {}


def tensors(name):
    if name not in _cache:
        wl = W.FULL[name]()
        _cache[name] = (wl, *wl.tensors())
    return _cache[name]


@pytest.mark.parametrize("name", ["c1", "c2"])
@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_full_size_materialised(ctx, name, planner):
    wl, a, b = tensors(name)
    print(f"full {wl.name} {planner}", flush=True)
    ref = _cache.get((name, "ref"))
    if ref is None:
        ref = oracle.contract_tiled(a, wl.a_modes, b, wl.b_modes, tile_size=128, high_precision=True)
        _cache[(name, "ref")] = ref
    bb = a if wl.self_contraction else b
    r = ctx.contract(a, list(wl.a_modes), bb, list(wl.b_modes), planner=planner)
    st = r.stats
    c, v = r.to_host()
    r.free()
    assert st["products"] == ref.products
    stats = assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"{wl.name} {planner}",
                               absum=ref.absum, nterms=ref.nterms)
    print(f"PARITY {wl.name} {planner} mirrored={st['mirrored']} {stats}", flush=True)


@pytest.mark.parametrize("name,stride", [("c3", 4096), ("c4", 32)])
@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_full_shape_sampled_tiles(ctx, name, stride, planner):
    """Full-shape operands, left rows x with (x / 128) % stride == 0."""
    wl, a, _ = tensors(name)
    print(f"sampled {wl.name} stride {stride} {planner}", flush=True)
    key = (name, "sref")
    if key not in _cache:
        _cache[key] = oracle.contract_tiled(a, wl.a_modes, a, wl.b_modes, tile_size=128, high_precision=True,
                                            sample_stride=stride)
    ref = _cache[key]
    r = ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), planner=planner, tile_hint=128,
                     left_tile_stride=stride)
    st = r.stats
    c, v = r.to_host()
    r.free()
    assert st["products"] == ref.products
    stats = assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS,
                               what=f"{wl.name} sampled/{stride} {planner}", absum=ref.absum, nterms=ref.nterms)
    print(f"PARITY {wl.name} sampled/{stride} {planner} {stats}", flush=True)


@pytest.mark.parametrize("name", ["c2", "c3", "c4"])
def test_full_size_mirror_on_off(ctx, name):
    """Whole contraction at full size: mirrored (learned planner, the measured
    configuration) and every output computed directly (exact planner) agree."""
    wl, a, _ = tensors(name)
    sink = name == "c3"  # 3.3e10 outputs: count + checksum
    out = {}
    for planner, mirror in (("learned", True), ("exact", False)):
        print(f"whole {wl.name} {planner} mirror={mirror}", flush=True)
        r = ctx.contract(a, list(wl.a_modes), a, list(wl.b_modes), planner=planner, mirror=mirror, sink=sink,
                         summarize=True)
        assert r.stats["mirrored"] == (1 if mirror else 0)
        out[mirror] = (r.nnz, r.coord_hash, r.val_sum, r.val_wsum, r.stats["products"])
        r.free()
    (n1, h1, s1, w1, p1), (n0, h0, s0, w0, p0) = out[True], out[False]
    assert n1 == n0 and h1 == h0 and p1 == p0
    assert s1 == pytest.approx(s0, rel=1e-9, abs=1e-9)
    assert w1 == pytest.approx(w0, rel=1e-9, abs=1e-9)
    record_parity(f"{wl.name} whole mirror on/off", outputs=int(n1), val_sum_rel_diff=abs(s1 - s0) / max(abs(s0), 1e-300),
                  val_wsum_rel_diff=abs(w1 - w0) / max(abs(w0), 1e-300))


def test_batched_shape_learned_fused(ctx):
    wls = W.c5_pairs(n_pairs=48, scale=0.1)
    pairs = [w.tensors() for w in wls]
    rs = ctx.contract_batch(pairs, list(wls[0].a_modes), list(wls[0].b_modes), planner="learned", sink=True)
    assert all(r.stats["units"] == rs[0].stats["units"] for r in rs)  # one fused launch
    # the per-pair checksums linearise over the bounding-box extents of the batch
    ea = np.max(np.stack([p[0].extents() for p in pairs]), axis=0)
    eb = np.max(np.stack([p[1].extents() for p in pairs]), axis=0)
    for i, ((a, b), r) in enumerate(zip(pairs, rs)):
        ref = oracle.contract_tiled(a, wls[i].a_modes, b, wls[i].b_modes, tile_size=128, sink=True, a_extents=ea,
                                    b_extents=eb)
        assert r.nnz == ref.nnz, i
        assert r.coord_hash == ref.coord_hash, i
        assert r.val_sum == pytest.approx(ref.val_sum, rel=1e-10, abs=1e-10), i
        assert r.val_wsum == pytest.approx(ref.val_wsum, rel=1e-10, abs=1e-10), i
    record_parity("c5 48 pairs at 10% learned fused", outputs=int(sum(r.nnz for r in rs)))
    for r in rs:
        r.free()
