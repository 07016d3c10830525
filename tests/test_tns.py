"""FROSTT text reader/writer (read.hpp:58-117 rules), Python and C++ header.

The reference's reader is pinned on hand-written files: comment and blank
lines skipped, the optional header line skipped, coordinates kept as
written, the last field the value; malformed lines rejected. The C++ drop-in
header's Tensor(filename, has_header) must read the same file identically."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT

from fastcc_amd.tns import read_tns, write_tns
from fastcc_amd.workloads import SparseTensor

SAMPLE = """# a FROSTT-style tensor
% another comment line
3 4 2
1 1 1 1.5
2 4 2 -0.25

3 2 1 1e-3
-1 1 1 9.0
"""


def _write(tmp, text):
    p = os.path.join(tmp, "t.tns")
    with open(p, "w") as f:
        f.write(text)
    return p


def test_read_rules():
    tmp = tempfile.mkdtemp(prefix="fcc_tns_")
    p = _write(tmp, SAMPLE)
    t = read_tns(p, has_header=True)
    # "-1 1 1 9.0" starts with '-': skipped like any non-digit line
    assert t.order == 3 and t.nnz == 3
    assert t.coords.tolist() == [[1, 2, 3], [1, 4, 2], [1, 2, 1]]
    assert t.vals.tolist() == [1.5, -0.25, 1e-3]
    # without has_header the shape line "3 4 2" is a nonzero (order 2, value 2)
    # and the order check rejects the following lines
    with pytest.raises(ValueError):
        read_tns(p, has_header=False)


@pytest.mark.parametrize("text", ["1  2 3.0\n", "1 2 3.0 \n", "1 x 3.0\n", "1 2 3.0\n1 2 3 4.0\n"])
def test_malformed_lines(text):
    tmp = tempfile.mkdtemp(prefix="fcc_tns_")
    with pytest.raises(ValueError):
        read_tns(_write(tmp, text))


def test_write_read_round_trip():
    rng = np.random.default_rng(11)
    coords = rng.integers(0, 50, size=(4, 200)).astype(np.uint32)
    vals = rng.standard_normal(200)
    t = SparseTensor(coords, vals)
    tmp = tempfile.mkdtemp(prefix="fcc_tns_")
    p = os.path.join(tmp, "rt.tns")
    write_tns(t, p, full_precision=True)
    with open(p) as f:
        first = f.readline()
    assert first == "".join(f"{int(x)} " for x in coords.max(axis=1)) + "\n"
    r = read_tns(p, has_header=True)
    assert np.array_equal(r.coords, coords) and np.array_equal(r.vals, vals)
    # default: C++ stream formatting, 6 significant digits
    p2 = os.path.join(tmp, "rt6.tns")
    write_tns(t, p2)
    r6 = read_tns(p2, has_header=True)
    assert np.array_equal(r6.coords, coords)
    assert np.allclose(r6.vals, vals, rtol=1e-5, atol=0) and not np.array_equal(r6.vals, vals)


MALFORMED = ["1  2 3.0\n", "1 2 3.0 \n", "1 x 3.0\n", "1 2 3.0\n1 2 3 4.0\n", "1 2 3.0x\n", "1 2\n3 4 5.0\n"]


@pytest.fixture(scope="module")
def cpp_reader(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("rd")
    src = tmp / "rd.cpp"
    src.write_text(r'''
#include <cstdio>
#include "fastcc.hpp"
int main(int argc, char **argv) {
    try {
        Tensor<double> t(argv[1], argv[2][0] == '1');
        const auto &c = t.coords_soa();
        const auto &v = t.values();
        size_t n = v.size();
        std::printf("%d %zu\n", t.get_dimensionality(), n);
        for (size_t i = 0; i < n; i++) {
            for (int m = 0; m < t.get_dimensionality(); m++) std::printf("%u ", c[(size_t)m * n + i]);
            std::printf("%.17g\n", v[i]);
        }
    } catch (const fastcc_error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
''')
    exe = tmp / "rd"
    host = os.path.join(ROOT, "fastcc-benchmark_amd", "host")
    lib = os.path.join(ROOT, "fastcc-benchmark_amd")
    subprocess.run(["g++", "-std=c++20", "-I", os.path.join(ROOT, "include"), "-I", host, "-o", str(exe), str(src),
                    "-L", lib, "-lfastcc_hip", f"-Wl,-rpath,{lib}"], check=True)
    return str(exe)


def test_cpp_header_reads_the_same(cpp_reader):
    tmp = tempfile.mkdtemp(prefix="fcc_tns_")
    p = _write(tmp, SAMPLE)
    out = subprocess.run([cpp_reader, p, "1"], capture_output=True, text=True, check=True).stdout.split("\n")
    order, n = map(int, out[0].split())
    t = read_tns(p, has_header=True)
    assert order == t.order and n == t.nnz
    for i in range(n):
        f = out[1 + i].split()
        assert [int(x) for x in f[:-1]] == t.coords[:, i].tolist()
        assert float(f[-1]) == t.vals[i]


@pytest.mark.parametrize("text", MALFORMED)
def test_cpp_header_rejects_malformed_lines(cpp_reader, text):
    """Doubled and trailing spaces, non-numeric fields and a changing order:
    the C++ reader refuses them like the Python reader (and the reference's
    std::stoi / std::stod)."""
    tmp = tempfile.mkdtemp(prefix="fcc_tns_")
    p = _write(tmp, text)
    with pytest.raises(ValueError):
        read_tns(p)
    r = subprocess.run([cpp_reader, p, "0"], capture_output=True, text=True)
    assert r.returncode == 2, (text, r.stdout, r.stderr)
