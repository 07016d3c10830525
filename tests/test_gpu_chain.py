"""GPU: chaining contractions on the device (fcc_result_as_tensor).

A result becomes the operand of the next contraction without a host round
trip (the reference goes through a host Tensor, contract.hpp:1824-1832).
Small-integer values keep every sum exact, so the chained result must equal
the oracle's chain bit for bit."""
import os

import numpy as np
import pytest

from conftest import ROOT, assert_same_result, rand_tensor

from oracle import oracle
from fastcc_amd import FastccError
from fastcc_amd.workloads import SparseTensor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("planner", ["exact", "learned"])
def test_chain_on_device(gpu_ctx, planner):
    print(f"planner={planner}", flush=True)
    gpu_ctx.set_mlp_params(np.load(os.path.join(ROOT, "fastcc-benchmark_amd", "fastcc_amd", "mlp_params.npy")))
    rng = np.random.default_rng(6060)
    a = rand_tensor(rng, (30, 12, 25), 2500, "smallint")   # a(i, k, j)
    b = rand_tensor(rng, (12, 40), 300, "smallint")        # b(k, l)
    e = rand_tensor(rng, (40, 7), 150, "smallint")         # e(l, m)
    # C(i, j, l) = sum_k a(i,k,j) b(k,l); D(i, j, m) = sum_l C(i,j,l) e(l,m)
    rc = gpu_ctx.contract(a, [1], b, [0], planner=planner)
    cdev = rc.as_tensor()
    assert rc.coord_stride == rc.nnz
    rd = gpu_ctx.contract(cdev, [2], e, [0], planner=planner)
    c_ref = oracle.contract_bruteforce(a, (1,), b, (0,))
    d_ref = oracle.contract_bruteforce(SparseTensor(c_ref.coords, c_ref.vals), (2,), e, (0,))
    c, v = rd.to_host()
    assert_same_result(c, v, d_ref.coords, d_ref.vals, rel=0, abs_floor=0, what=f"chain {planner}")
    # the intermediate is unchanged by the compaction
    c1, v1 = rc.to_host()
    assert_same_result(c1, v1, c_ref.coords, c_ref.vals, rel=0, abs_floor=0, what="intermediate")
    rd.free()
    rc.free()


def test_chain_rejects_sink_results(gpu_ctx):
    rng = np.random.default_rng(6061)
    a = rand_tensor(rng, (8, 5), 20, "smallint")
    b = rand_tensor(rng, (5, 6), 15, "smallint")
    r = gpu_ctx.contract(a, [1], b, [0], sink=True)
    if r.nnz:
        with pytest.raises(FastccError):
            r.as_tensor()
    r.free()
