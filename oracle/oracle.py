"""ctypes loader of the CPU oracle (oracle/libfcc_oracle.so) -- TEST
INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py, never by the product path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libfcc_oracle.so")


class _OcTensor(C.Structure):
    _fields_ = [("order", C.c_int32), ("nnz", C.c_int64), ("coords", C.c_void_p), ("vals", C.c_void_p),
                ("extents", C.c_void_p)]


class _OcResult(C.Structure):
    _fields_ = [
        ("order", C.c_int32),
        ("nnz", C.c_int64),
        ("coords", C.POINTER(C.c_uint32)),
        ("vals", C.POINTER(C.c_double)),
        ("absum", C.POINTER(C.c_double)),
        ("nterms", C.POINTER(C.c_uint32)),
        ("coord_hash", C.c_uint64),
        ("val_sum", C.c_double),
        ("val_wsum", C.c_double),
        ("products", C.c_uint64),
        ("seconds", C.c_double),
        ("index_seconds", C.c_double),
        ("threads", C.c_int32),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB} (make -C oracle)")
        _lib = C.CDLL(LIB)
        _lib.oc_contract_tiled.argtypes = [C.POINTER(_OcTensor), C.POINTER(C.c_int32), C.POINTER(_OcTensor),
                                           C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, C.POINTER(_OcResult)]
        _lib.oc_contract_tiled.restype = C.c_int
        _lib.oc_contract_tiled_ex.argtypes = [C.POINTER(_OcTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                              C.POINTER(_OcTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                              C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                              C.c_int32, C.POINTER(_OcResult)]
        _lib.oc_contract_tiled_ex.restype = C.c_int
        _lib.oc_contract_bruteforce_ex.argtypes = [C.POINTER(_OcTensor), C.POINTER(C.c_int32),
                                                   C.POINTER(C.c_int32), C.POINTER(_OcTensor),
                                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                                   C.c_int32, C.POINTER(_OcResult)]
        _lib.oc_contract_bruteforce_ex.restype = C.c_int
        _lib.oc_contract_bruteforce.argtypes = [C.POINTER(_OcTensor), C.POINTER(C.c_int32), C.POINTER(_OcTensor),
                                                C.POINTER(C.c_int32), C.c_int32, C.POINTER(_OcResult)]
        _lib.oc_contract_bruteforce.restype = C.c_int
        _lib.oc_result_sort.argtypes = [C.POINTER(_OcResult)]
        _lib.oc_result_sort.restype = C.c_int
        _lib.oc_result_free.argtypes = [C.POINTER(_OcResult)]
        _lib.oc_result_free.restype = None
        _lib.oc_mix64.argtypes = [C.c_uint64]
        _lib.oc_mix64.restype = C.c_uint64
    return _lib


class OracleResult:
    def __init__(self, order, coords, vals, coord_hash, val_sum, val_wsum, products, seconds, threads, nnz,
                 absum=None, nterms=None, index_seconds=0.0):
        self.order, self.coords, self.vals = order, coords, vals
        self.absum, self.nterms = absum, nterms
        self.coord_hash, self.val_sum, self.val_wsum = coord_hash, val_sum, val_wsum
        self.products, self.seconds, self.threads, self.nnz = products, seconds, threads, nnz
        self.index_seconds = index_seconds


def _mk(t, keep, ext=None):
    coords = np.ascontiguousarray(t.coords, dtype=np.uint32)
    vals = np.ascontiguousarray(t.vals, dtype=np.float64)
    keep += [coords, vals]
    o = _OcTensor()
    o.order, o.nnz = coords.shape[0], vals.shape[0]
    o.coords = coords.ctypes.data if o.nnz else None
    o.vals = vals.ctypes.data if o.nnz else None
    if ext is not None:
        e = np.ascontiguousarray(ext, dtype=np.uint32)
        keep.append(e)
        o.extents = e.ctypes.data
    return o


def _take(lib, r: _OcResult, sort: bool) -> OracleResult:
    coords = vals = absum = nterms = None
    if r.coords:
        if sort:
            rc = lib.oc_result_sort(C.byref(r))
            if rc:
                raise RuntimeError(f"oc_result_sort failed {rc}")
        n = r.nnz
        coords = np.ctypeslib.as_array(r.coords, shape=(max(r.order, 1) * max(n, 1),))[: r.order * n].copy()
        coords = coords.reshape(r.order, n)
        vals = np.ctypeslib.as_array(r.vals, shape=(max(n, 1),))[:n].copy()
        if r.absum:
            absum = np.ctypeslib.as_array(r.absum, shape=(max(n, 1),))[:n].copy()
            nterms = np.ctypeslib.as_array(r.nterms, shape=(max(n, 1),))[:n].copy()
    out = OracleResult(r.order, coords, vals, r.coord_hash, r.val_sum, r.val_wsum, r.products, r.seconds,
                       r.threads, r.nnz, absum, nterms, r.index_seconds)
    lib.oc_result_free(C.byref(r))
    return out


def _ilist(l):
    l = list(l)
    return (C.c_int32 * max(len(l), 1))(*l), len(l)


def contract_tiled(a, a_modes, b, b_modes, tile_size=128, threads=0, high_precision=True, sample_stride=1,
                   sink=False, sort=True, a_batch=(), b_batch=(), a_extents=None,
                   b_extents=None) -> OracleResult:
    """Restatement of the reference fastcc_multiply (see fcc_oracle.c); with
    batch lists, of the batch-mode multiply."""
    lib = _load()
    keep = []
    ta, tb = _mk(a, keep, a_extents), _mk(b, keep, b_extents)
    am, r = _ilist(a_modes)
    bm, _ = _ilist(b_modes)
    ab, nb = _ilist(a_batch)
    bb, _ = _ilist(b_batch)
    res = _OcResult()
    rc = lib.oc_contract_tiled_ex(C.byref(ta), am, ab, C.byref(tb), bm, bb, r, nb, int(tile_size), int(threads),
                                  1 if high_precision else 0, int(sample_stride), 1 if sink else 0, C.byref(res))
    if rc:
        raise RuntimeError(f"oc_contract_tiled failed {rc}")
    return _take(lib, res, sort)


def contract_bruteforce(a, a_modes, b, b_modes, a_batch=(), b_batch=(), a_extents=None,
                        b_extents=None) -> OracleResult:
    lib = _load()
    keep = []
    ta, tb = _mk(a, keep, a_extents), _mk(b, keep, b_extents)
    am, r = _ilist(a_modes)
    bm, _ = _ilist(b_modes)
    ab, nb = _ilist(a_batch)
    bb, _ = _ilist(b_batch)
    res = _OcResult()
    rc = lib.oc_contract_bruteforce_ex(C.byref(ta), am, ab, C.byref(tb), bm, bb, r, nb, C.byref(res))
    if rc:
        raise RuntimeError(f"oc_contract_bruteforce failed {rc}")
    return _take(lib, res, sort=False)


def mix64(x: int) -> int:
    return int(_load().oc_mix64(int(x) & 0xFFFFFFFFFFFFFFFF))
