"""Chunking of fused batches (csrc/fcc_chunk.h, host C++): the pair index of
a fused launch multiplies the left span and the key span, so a chunk holds
only as many pairs as keep the left span below 2^32 and key + column bits
within 64. Built with g++ here and checked on the CPU."""
import os
import subprocess

import pytest

from conftest import ROOT

DRIVER = r'''
#include <cstdio>
#include <vector>
#include "fcc_chunk.h"
int main() {
    int n, dA, dB, r, maxp;
    while (std::scanf("%d %d %d %d %d", &n, &dA, &dB, &r, &maxp) == 5) {
        std::vector<int32_t> am(r), bm(r);
        for (auto &x : am) std::scanf("%d", &x);
        for (auto &x : bm) std::scanf("%d", &x);
        std::vector<uint32_t> ea((size_t)n * dA), eb((size_t)n * dB);
        std::vector<uint64_t> na(n), nb(n);
        for (int i = 0; i < n; i++) {
            for (int m = 0; m < dA; m++) std::scanf("%u", &ea[(size_t)i * dA + m]);
            for (int m = 0; m < dB; m++) std::scanf("%u", &eb[(size_t)i * dB + m]);
            unsigned long long x, y;
            std::scanf("%llu %llu", &x, &y);
            na[i] = x;
            nb[i] = y;
        }
        auto ends = fcc::batch_chunks(n, dA, dB, am.data(), bm.data(), r, ea.data(), eb.data(), na.data(), nb.data(),
                                      maxp);
        for (auto e : ends) std::printf("%u ", e);
        std::printf("\n");
    }
}
'''


@pytest.fixture(scope="module")
def chunker(tmp_path_factory):
    d = tmp_path_factory.mktemp("chunk")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    inc = os.path.join(ROOT, "fastcc-benchmark_amd", "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + inc, str(src), "-o", str(exe)], check=True)

    def run(n, dA, dB, am, bm, ext_a, ext_b, nnz_a=10, nnz_b=10, maxp=4096):
        lines = [f"{n} {dA} {dB} {len(am)} {maxp}", " ".join(map(str, am)) or "", " ".join(map(str, bm)) or ""]
        for i in range(n):
            ea = ext_a(i) if callable(ext_a) else ext_a
            eb = ext_b(i) if callable(ext_b) else ext_b
            lines.append(" ".join(map(str, ea)) + " " + " ".join(map(str, eb)) + f" {nnz_a} {nnz_b}")
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        return [int(x) for x in out.stdout.split()]
    return run


def test_batched_shape_splits_below_2_to_32(chunker):
    # the batched workload's pairs: A 1024x128x1024x64, modes {1,3}: left span 2^20
    # per pair, so at most 4095 pairs keep the fused left span below 2^32
    ends = chunker(5000, 4, 3, [1, 3], [1, 2], [1024, 128, 1024, 64], [2048, 128, 64])
    assert ends == [4095, 5000]


def test_pair_cap_and_single_chunk(chunker):
    assert chunker(10000, 2, 2, [0], [0], [8, 8], [8, 8], maxp=4096) == [4096, 8192, 10000]
    assert chunker(7, 2, 2, [0], [0], [8, 8], [8, 8]) == [7]


def test_key_and_column_bits(chunker):
    # key radix 2^20 per pair, column span 2^30: key bits + column bits <= 64
    # leaves 14 bits for the pair index -> 2^14 pairs, capped by 4096 here
    ends = chunker(20000, 2, 2, [0], [0], [1 << 20, 2], [1 << 20, 1 << 30], maxp=1 << 20)
    assert ends[0] == 1 << 14 and ends[-1] == 20000


def test_growing_extents_and_oversized_pair(chunker):
    # pair 3 alone exceeds the left span: it forms a chunk of its own
    ext = lambda i: [2, 70000, 70000] if i == 3 else [2, 4, 4]
    ends = chunker(6, 3, 2, [0], [0], ext, [2, 4])
    assert ends == [3, 4, 6]
    # extents that grow along the batch end a chunk once the span is reached:
    # pair i has a left span of (i + 1) * 2^20
    ends = chunker(100, 2, 2, [0], [0], lambda i: [2, (i + 1) << 20], [2, 4])
    for c0, c1 in zip([0] + ends[:-1], ends):
        assert (c1 - c0) * (c1 << 20) < 1 << 32
        if c1 < 100:
            assert (c1 - c0 + 1) * ((c1 + 1) << 20) >= 1 << 32
