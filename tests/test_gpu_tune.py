"""GPU: the autotuner picks a tile shape per density bucket, later calls use
it, and results under the tuned shape equal the oracle's."""
import numpy as np
import pytest

from conftest import assert_same_result

from oracle import oracle
from fastcc_amd import Context
from fastcc_amd import workloads as W

pytestmark = pytest.mark.gpu

REL, ABS = 1e-12, 1e-14


@pytest.mark.parametrize("mk", [lambda: W.c1(0.25), lambda: W.c4(0.02)])
def test_autotune_records_and_applies(mk):
    ctx = Context(0)
    wl = mk()
    a, b = wl.tensors()
    b = a if wl.self_contraction else b
    best = ctx.autotune(a, list(wl.a_modes), b, list(wl.b_modes), reps=1)
    print(wl.name, best, flush=True)
    assert best["row_tile"] in [1 << k for k in range(11)]
    assert best["ms"] > 0.0
    table = ctx.get_tune()
    assert any(e["bucket"] == best["bucket"] and e["row_tile"] == best["row_tile"] for e in table)
    r = ctx.contract(a, list(wl.a_modes), b, list(wl.b_modes))
    assert r.stats["row_tile"] == best["row_tile"]
    if best["col_window"]:
        assert r.stats["col_window"] == best["col_window"]
    ref = oracle.contract_tiled(a, wl.a_modes, b, wl.b_modes, tile_size=128)
    c, v = r.to_host()
    assert_same_result(c, v, ref.coords, ref.vals, rel=REL, abs_floor=ABS, what=f"tuned {wl.name}",
                       absum=ref.absum, nterms=ref.nterms)
    r.free()
    ctx.close()
