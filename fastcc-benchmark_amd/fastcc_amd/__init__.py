"""fastcc_amd -- FaSTCC sparse-tensor contraction on MI355X (gfx950).

Python host layer over the C ABI of ``libfastcc_hip.so`` (include/fcc.h).
``Tensor`` / ``ListTensor`` / ``Tensor.fastcc_multiply`` mirror the
reference's C++ interface (contract.hpp:25-35, 824-827; index_tensors.hpp:80-126)
so tests read like uses of the reference; ``Context.contract`` exposes every
option of the C ABI. All computation happens in the HIP library: when it is
missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_GHASH, FCC_ACC_LHASH, FCC_PLAN_EXACT, FCC_PLAN_LEARNED,
                   FCC_SINK_COUNT_CHECKSUM, FCC_SINK_MATERIALIZE)

__all__ = ["Context", "Result", "DeviceTensor", "Tensor", "ListTensor", "FastccError", "TileAccumulator",
           "FCC_ACC_AUTO", "FCC_ACC_DENSE", "FCC_ACC_LHASH", "FCC_ACC_GHASH", "FCC_PLAN_EXACT", "FCC_PLAN_LEARNED"]


class FastccError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fcc error {code}: {msg}")
        self.code = code


@dataclass
class DeviceTensor:
    """Coordinates/values already resident on the GPU (e.g. torch tensors)."""
    order: int
    nnz: int
    coords_ptr: int   # device pointer, uint32 SoA [order][nnz]
    vals_ptr: int     # device pointer, float64 [nnz]
    keepalive: object = None
    given_extents: tuple = None  # per-mode extents, or None: max+1 inferred

    @staticmethod
    def from_torch(coords, vals):
        """coords: int32/uint32 CUDA tensor [order, nnz] (contiguous), vals: float64 CUDA tensor [nnz]."""
        assert coords.is_cuda and vals.is_cuda and coords.is_contiguous() and vals.is_contiguous()
        assert coords.element_size() == 4 and vals.element_size() == 8
        return DeviceTensor(int(coords.shape[0]), int(vals.shape[0]), coords.data_ptr(), vals.data_ptr(),
                            (coords, vals))


def _as_fcc_tensor(t, keep):
    ft = _abi.FccTensor()
    if isinstance(t, DeviceTensor):
        ft.order, ft.nnz = t.order, t.nnz
        ft.coords, ft.vals = t.coords_ptr, t.vals_ptr
        ft.on_device = 1
        ft.extents = None
        if t.given_extents is not None:
            e = (C.c_uint32 * ft.order)(*[int(x) for x in t.given_extents])
            keep.append(e)
            ft.extents = C.cast(e, C.POINTER(C.c_uint32))
        return ft
    coords = np.ascontiguousarray(t.coords, dtype=np.uint32)
    vals = np.ascontiguousarray(t.vals, dtype=np.float64)
    keep += [coords, vals]
    ft.order = int(coords.shape[0])
    ft.nnz = int(vals.shape[0])
    ft.coords = coords.ctypes.data if ft.nnz else None
    ft.vals = vals.ctypes.data if ft.nnz else None
    ft.extents = None
    ft.on_device = 0
    ext = getattr(t, "given_extents", None)
    if ext is not None:
        e = (C.c_uint32 * ft.order)(*[int(x) for x in ext])
        keep.append(e)
        ft.extents = C.cast(e, C.POINTER(C.c_uint32))
    return ft


def _stats_dict(s: _abi.FccStats):
    return {name: getattr(s, name) for name, _ in _abi.FccStats._fields_}


class Result:
    """A contraction result; device memory stays library-owned until free()."""

    def __init__(self, ctx: "Context", raw: _abi.FccResult):
        self._ctx = ctx
        self._raw = raw
        self.order = raw.order
        self.nnz = raw.nnz
        self.extents = tuple(raw.extents[i] for i in range(raw.order))
        self.coord_hash = raw.coord_hash
        self.val_sum = raw.val_sum
        self.val_wsum = raw.val_wsum
        self.stats = _stats_dict(raw.stats)
        self.coord_stride = raw.coord_stride
        self.coords_ptr = raw.coords or 0
        self.vals_ptr = raw.vals or 0

    def to_host(self):
        """(coords uint32 [order, nnz], vals float64 [nnz])"""
        coords = np.empty((self.order, self.nnz), dtype=np.uint32)
        vals = np.empty(self.nnz, dtype=np.float64)
        if self.nnz:
            rc = self._ctx.lib.fcc_result_to_host(self._ctx.h, C.byref(self._raw), coords.ctypes.data,
                                                  vals.ctypes.data)
            self._ctx._check(rc)
        return coords, vals

    def as_tensor(self) -> "DeviceTensor":
        """fcc_result_as_tensor: this result as an operand of a later
        contraction, still on the device (valid until free())."""
        t = _abi.FccTensor()
        self._ctx._check(self._ctx.lib.fcc_result_as_tensor(self._ctx.h, C.byref(self._raw), C.byref(t)))
        self.coord_stride = self._raw.coord_stride
        self.coords_ptr = self._raw.coords or 0
        return DeviceTensor(t.order, t.nnz, t.coords or 0, t.vals or 0, self, tuple(self.extents))

    def free(self):
        if self._raw is not None and self._ctx.h:
            self._ctx.lib.fcc_result_free(self._ctx.h, C.byref(self._raw))
        self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _vals_arg(v, keep):
    """(pointer, on_device) for a float64 numpy array or a CUDA tensor."""
    if v is None:
        return None, 0
    if hasattr(v, "is_cuda") and v.is_cuda:
        assert v.is_contiguous() and v.element_size() == 8
        keep.append(v)
        return v.data_ptr(), 1
    a = np.ascontiguousarray(v, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data, 0


class Plan:
    """A contraction planned for one coordinate pattern (fcc_plan_create):
    execute() reruns only the numeric pass on new values, optionally as a
    replayed hipGraph; info() and tiles() answer the symbolic queries."""

    def __init__(self, ctx: "Context", h):
        self._ctx = ctx
        self.h = h

    def info(self) -> dict:
        i = _abi.FccPlanInfo()
        self._ctx._check(self._ctx.lib.fcc_plan_info_get(self.h, C.byref(i)))
        d = {name: getattr(i, name) for name, _ in _abi.FccPlanInfo._fields_ if name != "extents"}
        d["extents"] = tuple(i.extents[m] for m in range(i.order))
        return d

    def tiles(self):
        """(exact nnz, products, predicted nnz or -1) per active left tile."""
        n = int(self._ctx.lib.fcc_plan_tile_count(self.h))
        nnz = np.zeros(n, dtype=np.uint64)
        prod = np.zeros(n, dtype=np.uint64)
        pred = np.zeros(n, dtype=np.float32)
        if n:
            self._ctx._check(self._ctx.lib.fcc_plan_tiles(self.h, nnz.ctypes.data, prod.ctypes.data,
                                                          pred.ctypes.data, n))
        return nnz, prod, pred

    def execute(self, a_vals, b_vals=None, graph=False) -> "Result":
        """New values in the entry order given at planning; the result views
        the plan's output buffers until the next execute."""
        keep = []
        pa, da = _vals_arg(a_vals, keep)
        pb, db = _vals_arg(b_vals, keep)
        if pb is not None and da != db:
            raise FastccError(_abi.FCC_ERR_INVALID, "a_vals and b_vals must both be host or both device")
        raw = _abi.FccResult()
        rc = self._ctx.lib.fcc_plan_execute(self.h, pa, pb, da, 1 if graph else 0, C.byref(raw))
        self._ctx._check(rc)
        return Result(self._ctx, raw)

    def close(self):
        if getattr(self, "h", None) and self._ctx.h:
            self._ctx.lib.fcc_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan(Plan):
    """A fused batch planned for its pairs' coordinate patterns
    (fcc_plan_create_batch): execute() reruns the numeric pass of every pair
    on new values, optionally as one replayed hipGraph."""

    def __init__(self, ctx: "Context", h, n: int, a_nnz, b_nnz):
        super().__init__(ctx, h)
        self.n = n
        self.a_off = np.concatenate([[0], np.cumsum(a_nnz)]).astype(np.int64)
        self.b_off = np.concatenate([[0], np.cumsum(b_nnz)]).astype(np.int64)

    def execute(self, a_vals, b_vals, graph=False) -> list:
        """a_vals / b_vals: every pair's values in pair order (one array each:
        host float64, or device tensors); one Result view per pair, valid
        until the next execute."""
        keep = []
        pa, da = _vals_arg(a_vals, keep)
        pb, db = _vals_arg(b_vals, keep)
        if da != db:
            raise FastccError(_abi.FCC_ERR_INVALID, "a_vals and b_vals must both be host or both device")
        raws = (_abi.FccResult * max(self.n, 1))()
        rc = self._ctx.lib.fcc_plan_execute_batch(self.h, pa, pb, da, 1 if graph else 0, raws)
        self._ctx._check(rc)
        return [Result(self._ctx, raws[i]) for i in range(self.n)]


class Context:
    """One GPU, one HIP stream, a caching device pool and the MLP planner."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = C.c_void_p()
        rc = self.lib.fcc_create(C.byref(h), int(device))
        if rc != 0:
            raise FastccError(rc, "fcc_create failed (no usable HIP device?)")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.fcc_last_error(self.h)
            raise FastccError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fcc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self.lib.fcc_stream(self.h) or 0

    @staticmethod
    def options(planner="exact", tile_hint=0, sink=False, force_acc=FCC_ACC_AUTO, col_window=0, workgroups=0,
                collect_stats=False, left_tile_stride=1, summarize=False, dense_keys=0.0, mirror=True,
                judge=False):
        o = _abi.FccOptions()
        _abi.load().fcc_options_default(C.byref(o))
        o.planner = FCC_PLAN_LEARNED if planner == "learned" else FCC_PLAN_EXACT
        o.tile_hint = int(tile_hint)
        o.sink = FCC_SINK_COUNT_CHECKSUM if sink else FCC_SINK_MATERIALIZE
        o.force_acc = int(force_acc)
        o.col_window = int(col_window)
        o.workgroups = int(workgroups)
        o.collect_stats = 1 if collect_stats else 0
        o.left_tile_stride = int(left_tile_stride)
        o.summarize = 1 if summarize else 0
        # dense-key GEMM threshold: 0 default, < 0 off, > 0 the threshold
        o.dense_keys = float(dense_keys)
        # self-contractions: compute each unordered output pair once (default) or every output
        o.mirror = 0 if mirror else -1
        # learned planner with mirroring: judge predictions by a sampled exact count
        o.judge = 1 if judge else 0
        return o

    def contract(self, a, a_modes, b, b_modes, a_batch=(), b_batch=(), **opts) -> Result:
        """fastcc_multiply (contract.hpp:824-827); with batch lists, the
        batch-mode multiply (contract.hpp:285-367): outputs carry (batch
        coords, A's other externals, B's other externals)."""
        keep = []
        ta = _as_fcc_tensor(a, keep)
        tb = ta if b is a else _as_fcc_tensor(b, keep)
        r, nb = len(a_modes), len(a_batch)
        if len(b_modes) != r or len(b_batch) != nb:
            raise FastccError(_abi.FCC_ERR_INVALID, "mode lists differ in length")
        am = (C.c_int32 * max(r, 1))(*a_modes)
        bm = (C.c_int32 * max(r, 1))(*b_modes)
        o = self.options(**opts)
        raw = _abi.FccResult()
        if nb:
            ab = (C.c_int32 * nb)(*a_batch)
            bb = (C.c_int32 * nb)(*b_batch)
            rc = self.lib.fcc_contract_ex(self.h, C.byref(ta), am, ab, C.byref(tb), bm, bb, r, nb, C.byref(o),
                                          C.byref(raw))
        else:
            rc = self.lib.fcc_contract(self.h, C.byref(ta), am, C.byref(tb), bm, r, C.byref(o), C.byref(raw))
        self._check(rc)
        return Result(self, raw)

    def plan(self, a, a_modes, b, b_modes, a_batch=(), b_batch=(), **opts) -> Plan:
        """fcc_plan_create: plans (and performs once) the contraction."""
        keep = []
        ta = _as_fcc_tensor(a, keep)
        tb = ta if b is a else _as_fcc_tensor(b, keep)
        r, nb = len(a_modes), len(a_batch)
        if len(b_modes) != r or len(b_batch) != nb:
            raise FastccError(_abi.FCC_ERR_INVALID, "mode lists differ in length")
        am = (C.c_int32 * max(r, 1))(*a_modes)
        bm = (C.c_int32 * max(r, 1))(*b_modes)
        ab = (C.c_int32 * max(nb, 1))(*a_batch)
        bb = (C.c_int32 * max(nb, 1))(*b_batch)
        o = self.options(**opts)
        h = C.c_void_p()
        rc = self.lib.fcc_plan_create(self.h, C.byref(ta), am, ab, C.byref(tb), bm, bb, r, nb, C.byref(o),
                                      C.byref(h))
        self._check(rc)
        return Plan(self, h)

    def plan_batch(self, pairs, a_modes, b_modes, **opts) -> BatchPlan:
        """fcc_plan_create_batch: plans (and performs once) the fused batch."""
        keep = []
        n = len(pairs)
        ta = (_abi.FccTensor * max(n, 1))()
        tb = (_abi.FccTensor * max(n, 1))()
        for i, (a, b) in enumerate(pairs):
            ta[i] = _as_fcc_tensor(a, keep)
            tb[i] = _as_fcc_tensor(b, keep)
        r = len(a_modes)
        am = (C.c_int32 * max(r, 1))(*a_modes)
        bm = (C.c_int32 * max(r, 1))(*b_modes)
        o = self.options(**opts)
        h = C.c_void_p()
        rc = self.lib.fcc_plan_create_batch(self.h, n, ta, am, tb, bm, r, C.byref(o), C.byref(h))
        self._check(rc)
        return BatchPlan(self, h, n, [ta[i].nnz for i in range(n)], [tb[i].nnz for i in range(n)])

    def contract_batch(self, pairs, a_modes, b_modes, **opts):
        keep = []
        n = len(pairs)
        ta = (_abi.FccTensor * max(n, 1))()
        tb = (_abi.FccTensor * max(n, 1))()
        for i, (a, b) in enumerate(pairs):
            ta[i] = _as_fcc_tensor(a, keep)
            tb[i] = _as_fcc_tensor(b, keep)
        r = len(a_modes)
        am = (C.c_int32 * max(r, 1))(*a_modes)
        bm = (C.c_int32 * max(r, 1))(*b_modes)
        o = self.options(**opts)
        raws = (_abi.FccResult * max(n, 1))()
        rc = self.lib.fcc_contract_batch(self.h, n, ta, am, tb, bm, r, C.byref(o), raws)
        self._check(rc)
        return [Result(self, raws[i]) for i in range(n)]

    # ---- device memory
    def set_memory_budget(self, nbytes: int):
        """Cap on the device bytes the context's pool may hold (0: none)."""
        self._check(self.lib.fcc_set_memory_budget(self.h, int(nbytes)))

    def memory_stats(self) -> dict:
        held, used = C.c_uint64(), C.c_uint64()
        self._check(self.lib.fcc_memory_stats(self.h, C.byref(held), C.byref(used)))
        return {"held": int(held.value), "in_use": int(used.value)}

    # ---- learned planner
    @property
    def mlp_param_count(self) -> int:
        return int(self.lib.fcc_mlp_param_count())

    def set_mlp_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.mlp_param_count
        self._check(self.lib.fcc_mlp_set_params(self.h, p.ctypes.data))

    def mlp_params(self):
        p = np.empty(self.mlp_param_count, dtype=np.float32)
        self._check(self.lib.fcc_mlp_get_params(self.h, p.ctypes.data))
        return p

    def tile_stats(self):
        n = int(self.lib.fcc_tile_stats_count(self.h))
        f = np.empty((n, _abi.FCC_MLP_IN), dtype=np.float32)
        l = np.empty((n, _abi.FCC_MLP_OUT), dtype=np.float32)
        if n:
            self._check(self.lib.fcc_tile_stats_copy(self.h, f.ctypes.data, l.ctypes.data, n))
        return f, l

    def train_mlp(self, features, labels, epochs=30, batch=256, lr=3e-3, seed=1):
        f = np.ascontiguousarray(features, dtype=np.float32)
        l = np.ascontiguousarray(labels, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.fcc_mlp_train(self.h, f.ctypes.data, l.ctypes.data, f.shape[0], int(epochs),
                                           int(batch), float(lr), int(seed), C.byref(loss)))
        return float(loss.value)

    def mlp_forward_backward(self, features, labels):
        f = np.ascontiguousarray(features, dtype=np.float32)
        l = np.ascontiguousarray(labels, dtype=np.float32)
        n = f.shape[0]
        pred = np.empty((n, _abi.FCC_MLP_OUT), dtype=np.float32)
        grad = np.empty(self.mlp_param_count, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.fcc_mlp_forward_backward(self.h, f.ctypes.data, l.ctypes.data, n, pred.ctypes.data,
                                                      C.byref(loss), grad.ctypes.data))
        return pred, float(loss.value), grad

    # ---- autotuner
    def autotune(self, a, a_modes, b, b_modes, reps=2, **opts) -> dict:
        """fcc_autotune: time tile shapes on this contraction and record the
        fastest for its density bucket in the context's table."""
        keep = []
        ta = _as_fcc_tensor(a, keep)
        tb = ta if b is a else _as_fcc_tensor(b, keep)
        r = len(a_modes)
        am = (C.c_int32 * max(r, 1))(*a_modes)
        bm = (C.c_int32 * max(r, 1))(*b_modes)
        o = self.options(**opts)
        e = _abi.FccTuneEntry()
        self._check(self.lib.fcc_autotune(self.h, C.byref(ta), am, C.byref(tb), bm, r, C.byref(o), int(reps),
                                          C.byref(e)))
        return {"bucket": e.bucket, "row_tile": e.row_tile, "col_window": e.col_window,
                "workgroups": e.workgroups, "ms": e.ms}

    # ---- autotuner table
    def set_tune(self, entries):
        arr = (_abi.FccTuneEntry * max(len(entries), 1))()
        for i, e in enumerate(entries):
            arr[i].bucket, arr[i].row_tile, arr[i].col_window, arr[i].workgroups = (
                int(e["bucket"]), int(e["row_tile"]), int(e["col_window"]), int(e.get("workgroups", 0)))
            arr[i].ms = float(e.get("ms", 0.0))
        self._check(self.lib.fcc_tune_set(self.h, arr, len(entries)))

    def get_tune(self):
        arr = (_abi.FccTuneEntry * 64)()
        n = C.c_int32()
        self._check(self.lib.fcc_tune_get(self.h, arr, 64, C.byref(n)))
        return [dict(bucket=arr[i].bucket, row_tile=arr[i].row_tile, col_window=arr[i].col_window,
                     workgroups=arr[i].workgroups, ms=arr[i].ms) for i in range(n.value)]

    def density_bucket(self, d: float) -> int:
        return int(self.lib.fcc_density_bucket(float(d)))


# ---------------------------------------------------------------------------
# Mirror of the reference interface (contract.hpp / index_tensors.hpp).

class TileAccumulator:
    """Tag standing for TileAccumulator<double> (index_tensors.hpp:994-1065):
    the planner's dense LDS tile; hash kinds cover TileAccumulatorMap."""


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class ListTensor:
    """Result of fastcc_multiply (index_tensors.hpp:80-126). The reference
    exposes only compute_nnz_count(); coords()/values() are additions."""

    def __init__(self, result: Result):
        self._r = result
        self._host = None

    def compute_nnz_count(self) -> int:
        return int(self._r.nnz)

    def _fetch(self):
        if self._host is None:
            self._host = self._r.to_host()
        return self._host

    def coords(self) -> np.ndarray:
        return self._fetch()[0]

    def values(self) -> np.ndarray:
        return self._fetch()[1]

    @property
    def result(self) -> Result:
        return self._r


class Tensor:
    """Sparse tensor in COO form (contract.hpp:25-35). ``coords`` is
    [order, nnz] uint32, ``vals`` [nnz] float64."""

    def __init__(self, coords, vals):
        self.coords = np.ascontiguousarray(coords, dtype=np.uint32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.dimensionality = 42
        self.shape = None

    # contract.hpp:101-105
    def _infer_dimensionality(self):
        if self.vals.shape[0] > 0:
            self.dimensionality = int(self.coords.shape[0])

    # contract.hpp:124-144 (shape = max coordinate per mode)
    def _infer_shape(self):
        if self.vals.shape[0] > 0:
            self.shape = self.coords.max(axis=1).astype(np.int64)

    def get_dimensionality(self):
        if self.dimensionality == 42:
            self._infer_dimensionality()
        return self.dimensionality

    def get_shape_ref(self):
        if self.shape is None:
            self._infer_dimensionality()
            self._infer_shape()
        return self.shape

    def get_size(self):
        return int(self.vals.shape[0])

    def fastcc_multiply(self, other: "Tensor", left_contr, right_contr, tile_size=100, *, ctx: Context | None = None,
                        **opts) -> ListTensor:
        """contract.hpp:824-827: tile_size rounds up to a power of two and sets
        the rows per left tile."""
        ctx = ctx or default_context()
        r = ctx.contract(self, list(left_contr), other, list(right_contr), tile_hint=tile_size, **opts)
        return ListTensor(r)
