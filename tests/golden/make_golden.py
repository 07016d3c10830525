"""Writes tests/golden/oracle_small.json: tiny contraction cases (inputs and
the oracle's outputs, small-integer values so every sum is exact).

The reference has no tests or golden vectors of its own and cannot be built
here, so these vectors are regression fixtures of the oracle; the test that
reads them also re-derives every case with numpy einsum.
Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "fastcc-benchmark_amd"), ROOT, os.path.join(ROOT, "tests")]

from conftest import rand_tensor  # noqa: E402
from oracle import oracle  # noqa: E402

CASES = [
    ("readme_shape", (4, 3, 5, 2), 18, (1, 3), (3, 3, 2), 10, (1, 2), 2),
    ("self_two_modes", (6, 3, 4, 2), 30, (0, 1), None, None, (0, 1), 4),
    ("nonleading_pairs", (3, 4, 5), 20, (2, 0), (5, 2, 3), 12, (0, 2), 1),
    ("outer", (3, 2), 4, (), (2, 2), 3, (), 8),
]


def main():
    rng = np.random.default_rng(20261203)
    out = []
    for name, da, na, am, db, nb, bm, tile in CASES:
        a = rand_tensor(rng, da, na, "smallint")
        b = a if db is None else rand_tensor(rng, db, nb, "smallint")
        r = oracle.contract_tiled(a, am, b, bm, tile_size=tile, threads=1)
        out.append(dict(name=name, a_coords=a.coords.tolist(), a_vals=a.vals.tolist(), a_modes=list(am),
                        b_coords=b.coords.tolist(), b_vals=b.vals.tolist(), b_modes=list(bm), tile=tile,
                        c_coords=r.coords.tolist(), c_vals=r.vals.tolist()))
    with open(os.path.join(ROOT, "tests", "golden", "oracle_small.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
