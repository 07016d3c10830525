"""Register budget of the accumulate kernels (compiled here for gfx950).

k_numeric / k_count keep their LDS accumulator in about 78 KB, so two
512-thread workgroups share a CU only when each wavefront needs at most 128
VGPRs (4 waves per SIMD). One register too many halves the resident
workgroups and, on the hash-heavy count + checksum workloads, roughly halves
throughput; this test fails on that before any GPU run. No variant may spill
to scratch either."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_accumulate_kernels_fit_two_workgroups_per_cu(tmp_path):
    pkg = os.path.join(ROOT, "fastcc-benchmark_amd")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           "-c", os.path.join(pkg, "csrc", "fcc_accum.hip"), "-o", str(tmp_path / "acc.o"),
           "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    kernels = {}
    name = None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        for key, pat in (("vgpr", r"VGPRs: (\d+)"), ("scratch", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("occ", r"Occupancy \[waves/SIMD\]: (\d+)")):
            m = re.search(pat, line)
            if m and name:
                kernels[name][key] = int(m.group(1))
    acc = {k: v for k, v in kernels.items() if "k_numeric" in k or "k_count" in k}
    assert len(acc) >= 9, sorted(kernels)
    for k, v in acc.items():
        assert v.get("scratch", 0) == 0, f"{k} spills {v['scratch']} bytes per lane"
        assert v.get("vgpr", 999) <= 128 and v.get("occ", 0) >= 4, f"{k}: {v}"
