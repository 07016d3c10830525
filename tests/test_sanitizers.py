"""Address and undefined-behaviour sanitizer runs of the CPU-side code (host
compilers only; no GPU code is built here):

* the oracle (oracle/fcc_oracle.c) with gcc -fsanitize=address,undefined,
  driven over random small contractions: the tiled restatement against the
  brute force (bit-exact with small-integer values), count + checksum sinks,
  sampled left tiles, batch pairs, empty operands, 1-4 OpenMP threads;
* the host C++ of the drop-in header (host/fastcc.hpp: CoOrdinate, the FROSTT
  reader on well-formed and malformed files) and the fused-batch chunk planner
  (csrc/fcc_chunk.h) with g++ -fsanitize=address,undefined.
Any sanitizer report makes the run exit non-zero."""
import os
import subprocess

import pytest

from conftest import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")

ORACLE_DRIVER = r'''
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fcc_oracle.h"

static uint64_t rs = 88172645463325252ull;
static uint64_t rnd(void) { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }

/* unique random coordinates, small-integer values */
static void make(oc_tensor *t, int order, const uint32_t *dims, int64_t want, uint32_t **c, double **v) {
    uint64_t space = 1;
    for (int m = 0; m < order; m++) space *= dims[m];
    if ((uint64_t)want > space) want = (int64_t)space;
    unsigned char *used = calloc(space, 1);
    *c = malloc(sizeof(uint32_t) * order * (want ? want : 1));
    *v = malloc(sizeof(double) * (want ? want : 1));
    int64_t n = 0;
    while (n < want) {
        uint64_t lin = rnd() % space;
        if (used[lin]) continue;
        used[lin] = 1;
        for (int m = order - 1; m >= 0; m--) { (*c)[(size_t)m * want + n] = (uint32_t)(lin % dims[m]); lin /= dims[m]; }
        (*v)[n] = (double)((int)(rnd() % 7) - 3);
        n++;
    }
    free(used);
    t->order = order; t->nnz = want; t->coords = *c; t->vals = *v; t->extents = NULL;
}

static int same(oc_result *x, oc_result *y) {
    if (x->nnz != y->nnz || x->order != y->order) return 0;
    if (x->nnz && (memcmp(x->coords, y->coords, sizeof(uint32_t) * x->order * x->nnz) ||
                   memcmp(x->vals, y->vals, sizeof(double) * x->nnz))) return 0;
    return 1;
}

int main(void) {
    int fails = 0;
    for (int it = 0; it < 40; it++) {
        uint32_t da[3] = {2 + rnd() % 9, 2 + rnd() % 6, 2 + rnd() % 9}, db[2] = {da[1], 2 + rnd() % 9};
        oc_tensor A, B;
        uint32_t *ca, *cb;
        double *va, *vb;
        make(&A, 3, da, it % 7 == 0 ? 0 : (int64_t)(rnd() % 200), &ca, &va);
        make(&B, 2, db, (int64_t)(rnd() % 60), &cb, &vb);
        int32_t am[1] = {1}, bm[1] = {0};
        oc_result bf, tl, sk, sm;
        memset(&bf, 0, sizeof bf); memset(&tl, 0, sizeof tl); memset(&sk, 0, sizeof sk); memset(&sm, 0, sizeof sm);
        int T = 1 << (rnd() % 6), th = 1 + it % 4;
        if (oc_contract_bruteforce(&A, am, &B, bm, 1, &bf)) fails++;
        if (oc_contract_tiled(&A, am, &B, bm, 1, T, th, 1, 1, 0, &tl) || oc_result_sort(&tl)) fails++;
        if (!same(&bf, &tl)) { printf("tiled != brute force at %d\n", it); fails++; }
        if (oc_contract_tiled(&A, am, &B, bm, 1, T, th, 0, 1, 1, &sk)) fails++;
        if (sk.nnz != bf.nnz || sk.coord_hash != tl.coord_hash) { printf("sink differs at %d\n", it); fails++; }
        if (oc_contract_tiled(&A, am, &B, bm, 1, T, th, 0, 3, 0, &sm)) fails++;
        /* batch pair (mode 0 of A with mode 1 of B) and self-contraction */
        int32_t ab[1] = {0}, bb[1] = {1};
        oc_result x, y, z;
        memset(&x, 0, sizeof x); memset(&y, 0, sizeof y); memset(&z, 0, sizeof z);
        if (oc_contract_tiled_ex(&A, am, ab, &B, bm, bb, 1, 1, T, th, 1, 1, 0, &x) || oc_result_sort(&x)) fails++;
        if (oc_contract_bruteforce_ex(&A, am, ab, &B, bm, bb, 1, 1, &y)) fails++;
        if (!same(&x, &y)) { printf("batch tiled != brute force at %d\n", it); fails++; }
        int32_t sm2[2] = {0, 1};
        if (oc_contract_tiled(&A, sm2, &A, sm2, 2, T, th, 1, 1, 0, &z)) fails++;
        oc_result_free(&bf); oc_result_free(&tl); oc_result_free(&sk); oc_result_free(&sm);
        oc_result_free(&x); oc_result_free(&y); oc_result_free(&z);
        free(ca); free(va); free(cb); free(vb);
    }
    printf("fails %d\n", fails);
    return fails ? 1 : 0;
}
'''

HOST_DRIVER = r'''
#include <cstdio>
#include <fstream>
#include "fastcc.hpp"
#include "fcc_chunk.h"
int main(int argc, char **argv) {
    int bad = 0;
    // CoOrdinate brace-initialisation and the reader on each file given
    CoOrdinate a{1, 3}, b{1, 2};
    if (a.get_dimensionality() != 2 || b.get_index(1) != 2) bad++;
    for (int i = 1; i < argc; i++) {
        try {
            Tensor<double> t(argv[i], false);
            t._infer_shape();
            std::printf("%s: %d nnz\n", argv[i], t.get_size());
        } catch (const fastcc_error &e) {
            std::printf("%s: rejected\n", argv[i]);
        }
    }
    // chunk planner on the batched shape and on growing extents
    std::vector<uint32_t> ea, eb;
    std::vector<uint64_t> na, nb;
    for (int i = 0; i < 5000; i++) {
        uint32_t e[4] = {1024, 128, 1024, 64}, f[3] = {2048, 128, 64};
        ea.insert(ea.end(), e, e + 4);
        eb.insert(eb.end(), f, f + 3);
        na.push_back(1000 + i);
        nb.push_back(10 + i % 7);
    }
    int32_t am[2] = {1, 3}, bm[2] = {1, 2};
    auto ends = fcc::batch_chunks(5000, 4, 3, am, bm, 2, ea.data(), eb.data(), na.data(), nb.data(), 4096);
    if (ends.size() != 2 || ends[0] != 4095 || ends[1] != 5000) bad++;
    std::printf("bad %d\n", bad);
    return bad;
}
'''


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=ENV, **kw)


def test_oracle_under_asan_ubsan(tmp_path):
    src = tmp_path / "drv.c"
    src.write_text(ORACLE_DRIVER)
    exe = tmp_path / "drv"
    odir = os.path.join(ROOT, "oracle")
    b = _run(["gcc", *SAN, "-fopenmp", "-std=gnu11", "-I", odir, str(src), os.path.join(odir, "fcc_oracle.c"),
              "-lm", "-o", str(exe)])
    assert b.returncode == 0, b.stderr[-3000:]
    r = _run([str(exe)])
    assert r.returncode == 0 and "fails 0" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_host_cpp_under_asan_ubsan(tmp_path):
    src = tmp_path / "host.cpp"
    src.write_text(HOST_DRIVER)
    exe = tmp_path / "host"
    pkg = os.path.join(ROOT, "fastcc-benchmark_amd")
    # the header's reader and types are host code; the HIP library itself is
    # not linked (no entry point is called)
    b = _run(["g++", *SAN, "-std=c++20", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(pkg, "host"),
              "-I", os.path.join(pkg, "csrc"), str(src), "-o", str(exe), "-Wl,--unresolved-symbols=ignore-all"])
    assert b.returncode == 0, b.stderr[-3000:]
    files = {"good.tns": "# c\n1 2 3 1.5\n2 2 2 -2\n", "dbl.tns": "1  2 3.0\n", "trail.tns": "1 2 3.0 \n",
             "alpha.tns": "1 x 3.0\n", "order.tns": "1 2 3.0\n1 2 3 4.0\n", "empty.tns": ""}
    paths = []
    for name, text in files.items():
        p = tmp_path / name
        p.write_text(text)
        paths.append(str(p))
    r = _run([str(exe), *paths])
    assert r.returncode == 0 and "bad 0" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "good.tns: 2 nnz" in r.stdout
    for name in ("dbl", "trail", "alpha", "order"):
        assert f"{name}.tns: rejected" in r.stdout, r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
