"""CPU checks of the boundary: the HIP library loads, exports every entry
point include/fcc.h declares, and the C++ drop-in header compiles. No kernel
is launched here."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

from fastcc_amd import _abi

HEADER = os.path.join(ROOT, "include", "fcc.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fcc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 15
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"
    assert sorted(_abi.EXPORTS) == decl


def test_nm_lists_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    for s in declared_symbols():
        assert re.search(rf"\bT {s}\b", out), s


def test_library_has_gfx950_code_object():
    data = open(_abi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_pure_host_entry_points():
    lib = _abi.load()
    assert lib.fcc_mlp_param_count() == 6 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2 + 2 * 6 + 2 * 2
    assert lib.fcc_density_bucket(0.0) == 0
    assert lib.fcc_density_bucket(1e-3) == 3
    assert lib.fcc_density_bucket(1.0) == 6
    assert lib.fcc_density_bucket(500.0) == 8
    o = _abi.FccOptions()
    lib.fcc_options_default(ctypes.byref(o))
    assert o.planner == 0 and o.sink == 0 and o.left_tile_stride == 1
    # a NULL context is rejected without touching a device
    assert lib.fcc_contract(None, None, None, None, None, 0, None, None) == _abi.FCC_ERR_INVALID
    assert lib.fcc_tune_set(None, None, 0) == _abi.FCC_ERR_INVALID


def test_struct_layout_matches_header():
    """Compile a tiny C program printing sizeof/offsetof of the ABI structs and
    compare with the ctypes mirrors."""
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "fcc.h"
int main(void){
 printf("%zu %zu %zu %zu %zu\n", sizeof(fcc_tensor), sizeof(fcc_options), sizeof(fcc_stats), sizeof(fcc_result), sizeof(fcc_tune_entry));
 printf("%zu %zu %zu\n", offsetof(fcc_result, stats), offsetof(fcc_result, coord_stride), offsetof(fcc_stats, ms_done));
 return 0;}
'''
    import tempfile
    tmp = os.path.join(tempfile.mkdtemp(prefix="fcc_abi_"), "abi_layout")
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", tmp, tmp + ".c"], check=True)
    line1, line2 = subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    sizes = [int(x) for x in line1.split()]
    assert sizes == [ctypes.sizeof(_abi.FccTensor), ctypes.sizeof(_abi.FccOptions), ctypes.sizeof(_abi.FccStats),
                     ctypes.sizeof(_abi.FccResult), ctypes.sizeof(_abi.FccTuneEntry)]
    offs = [int(x) for x in line2.split()]
    assert offs == [_abi.FccResult.stats.offset, _abi.FccResult.coord_stride.offset,
                    _abi.FccStats.ms_done.offset]


def test_cpp_dropin_header_compiles():
    host = os.path.join(ROOT, "fastcc-benchmark_amd", "host")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-I", host,
                        os.path.join(host, "readme_example.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
