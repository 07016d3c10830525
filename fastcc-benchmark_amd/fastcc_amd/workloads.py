"""Synthetic contraction workloads (seeded, deterministic).

Each workload is a pair of sparse tensors plus the paired contraction mode
lists, generated by the C generator in ``csrc/fcc_gen.c``.  The five full-size
shapes are the measured configurations; ``scaled`` variants keep the same
structure (orders, mode lists, Zipf exponents, value laws) at sizes the CPU
oracle finishes in seconds.

The reference itself reads FROSTT files (read.hpp:58-106) and always
self-contracts one tensor on one mode list (benchmark.cpp:60-70); shapes here
that use two different tensors follow the README usage (README.md:8-13).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)

UNIFORM_PM1, UNIFORM_01, LOGNORMAL, NORMAL, ONES, SMALLINT = range(6)
MAX_ORDER = 8
SEED_BASE = 0xFA57CC00


class _Spec(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int32),
        ("dims", ctypes.c_uint32 * MAX_ORDER),
        ("zipf_s", ctypes.c_double * MAX_ORDER),
        ("nnz", ctypes.c_int64),
        ("value_dist", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


_lib = None


def _gen_lib():
    global _lib
    if _lib is None:
        path = os.path.join(_PKG, "libfcc_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"generator library missing: {path} (run __graft_entry__.build())")
        _lib = ctypes.CDLL(path)
        _lib.fccgen_generate.restype = ctypes.c_int
        _lib.fccgen_generate.argtypes = [ctypes.POINTER(_Spec), ctypes.c_void_p, ctypes.c_void_p]
    return _lib


@dataclass
class TensorSpec:
    dims: tuple
    nnz: int
    zipf: dict = field(default_factory=dict)  # mode -> s
    value_dist: int = UNIFORM_PM1
    seed: int = SEED_BASE

    @property
    def order(self):
        return len(self.dims)


@dataclass
class SparseTensor:
    """Host COO tensor, SoA coordinates: coords[m, n]."""
    coords: np.ndarray  # uint32 [order, nnz]
    vals: np.ndarray    # float64 [nnz]

    @property
    def order(self):
        return int(self.coords.shape[0])

    @property
    def nnz(self):
        return int(self.vals.shape[0])

    def extents(self):
        if self.nnz == 0:
            return np.zeros(self.order, dtype=np.uint64)
        return self.coords.max(axis=1).astype(np.uint64) + 1


def generate(spec: TensorSpec) -> SparseTensor:
    lib = _gen_lib()
    s = _Spec()
    s.order = spec.order
    for m, d in enumerate(spec.dims):
        s.dims[m] = int(d)
        s.zipf_s[m] = float(spec.zipf.get(m, 0.0))
    s.nnz = int(spec.nnz)
    s.value_dist = int(spec.value_dist)
    s.seed = int(spec.seed) & 0xFFFFFFFFFFFFFFFF
    coords = np.empty((spec.order, spec.nnz), dtype=np.uint32)
    vals = np.empty(spec.nnz, dtype=np.float64)
    rc = lib.fccgen_generate(ctypes.byref(s), coords.ctypes.data, vals.ctypes.data)
    if rc != 0:
        raise ValueError(f"fccgen_generate failed ({rc}) for {spec}")
    return SparseTensor(coords, vals)


@dataclass
class Workload:
    name: str
    a: TensorSpec
    a_modes: tuple
    b: TensorSpec | None      # None: self-contraction of a with itself
    b_modes: tuple
    description: str = ""

    @property
    def self_contraction(self):
        return self.b is None

    def tensors(self):
        ta = generate(self.a)
        tb = ta if self.b is None else generate(self.b)
        return ta, tb


def _seed(cfg, k):
    return SEED_BASE + 16 * cfg + k


def c1(scale=1.0):
    """order-4 x order-3, uniform coordinates, result(i,j,k)=A(i,c0,j,c1)*B(k,c0,c1)."""
    s = scale
    a = TensorSpec((max(2, int(500 * s)), 64, max(2, int(500 * s)), 32), int(200_000 * s * s),
                   value_dist=UNIFORM_PM1, seed=_seed(1, 0))
    b = TensorSpec((max(2, int(300 * s)), 64, 32), int(100_000 * s), value_dist=UNIFORM_PM1, seed=_seed(1, 1))
    return Workload("c1_order4x3_uniform" + ("" if s == 1 else f"_s{s:g}"), a, (1, 3), b, (1, 2),
                    "A 500x64x500x32 200k nnz x B 300x64x32 100k nnz, modes {1,3}x{1,2}")


def c2(scale=1.0):
    """order-4 self-contraction, mode-0 Zipf(1.1), lognormal values, dense output."""
    s = scale
    t = TensorSpec((max(2, int(6186 * s)), 24, max(2, int(77 * math.sqrt(s))), max(2, int(32 * math.sqrt(s)))),
                   int(5_300_000 * s * s), zipf={0: 1.1}, value_dist=LOGNORMAL, seed=_seed(2, 0))
    return Workload("c2_order4_self_zipf1.1" + ("" if s == 1 else f"_s{s:g}"), t, (0, 1), None, (0, 1),
                    "T 6186x24x77x32 5.3M nnz, modes {0,1} with itself")


def c3(scale=1.0):
    """order-4 self-contraction, mode-2 Zipf(1.2), hash-heavy sparse output."""
    s = scale
    t = TensorSpec((max(2, int(2482 * s)), max(2, int(2862 * s)), max(2, int(14036 * s)), 17),
                   int(3_100_000 * s), zipf={2: 1.2}, value_dist=UNIFORM_01, seed=_seed(3, 0))
    return Workload("c3_order4_self_zipf1.2" + ("" if s == 1 else f"_s{s:g}"), t, (2, 3), None, (2, 3),
                    "T 2482x2862x14036x17 3.1M nnz, modes {2,3} with itself")


def c4(scale=1.0):
    """order-5 self-contraction, modes 0,1 Zipf(1.05), order-6 output."""
    s = scale
    t = TensorSpec((max(2, int(165427 * s)), max(2, int(11374 * s)), 2, max(2, int(100 * math.sqrt(s))),
                    max(2, int(89 * math.sqrt(s)))),
                   int(26_000_000 * s), zipf={0: 1.05, 1: 1.05}, value_dist=UNIFORM_PM1, seed=_seed(4, 0))
    return Workload("c4_order5_self_zipf1.05" + ("" if s == 1 else f"_s{s:g}"), t, (0, 1), None, (0, 1),
                    "T 165427x11374x2x100x89 26M nnz, modes {0,1} with itself")


def c5_pairs(n_pairs=256, scale=1.0):
    """Uneven batch: A_b 1024x128x1024x64 x B_b 2048x128x64, modes {1,3}x{1,2},
    nnz log-uniform (A 10k-2M, B 5k-1M), mode-1 Zipf(1.1), normal values."""
    rng = np.random.Generator(np.random.PCG64(_seed(5, 0)))
    out = []
    for p in range(n_pairs):
        na = int(math.floor(math.exp(rng.uniform(math.log(1e4), math.log(2e6))) * scale))
        nb = int(math.floor(math.exp(rng.uniform(math.log(5e3), math.log(1e6))) * scale))
        a = TensorSpec((1024, 128, 1024, 64), max(1, na), zipf={1: 1.1}, value_dist=NORMAL,
                       seed=_seed(5, 1) + 2 * p * 7919)
        b = TensorSpec((2048, 128, 64), max(1, nb), zipf={1: 1.1}, value_dist=NORMAL,
                       seed=_seed(5, 2) + 2 * p * 7919 + 1)
        out.append(Workload(f"c5_pair{p}", a, (1, 3), b, (1, 2)))
    return out


FULL = {"c1": c1, "c2": c2, "c3": c3, "c4": c4}
