"""ctypes bindings of include/fcc.h (the C ABI of libfastcc_hip.so)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_PKG, "libfastcc_hip.so")

FCC_MAX_ORDER = 8
FCC_OK, FCC_ERR_INVALID, FCC_ERR_OOM, FCC_ERR_OVERFLOW, FCC_ERR_HIP = 0, -1, -2, -3, -4
FCC_PLAN_EXACT, FCC_PLAN_LEARNED = 0, 1
FCC_SINK_MATERIALIZE, FCC_SINK_COUNT_CHECKSUM = 0, 1
FCC_ACC_AUTO, FCC_ACC_DENSE, FCC_ACC_LHASH, FCC_ACC_GHASH = 0, 1, 2, 3
FCC_MLP_IN, FCC_MLP_HID, FCC_MLP_OUT = 6, 64, 2

# every symbol include/fcc.h declares
EXPORTS = (
    "fcc_create", "fcc_destroy", "fcc_last_error", "fcc_stream", "fcc_options_default", "fcc_contract",
    "fcc_contract_ex", "fcc_plan_create", "fcc_plan_info_get", "fcc_plan_tile_count", "fcc_plan_tiles",
    "fcc_plan_execute", "fcc_plan_destroy", "fcc_autotune", "fcc_result_as_tensor",
    "fcc_contract_batch", "fcc_result_to_host", "fcc_result_free", "fcc_mlp_param_count", "fcc_mlp_set_params",
    "fcc_mlp_get_params", "fcc_tile_stats_count", "fcc_tile_stats_copy", "fcc_mlp_train",
    "fcc_mlp_forward_backward", "fcc_density_bucket", "fcc_tune_set", "fcc_tune_get",
    "fcc_set_memory_budget", "fcc_memory_stats", "fcc_contract_noalloc", "fcc_plan_create_batch",
    "fcc_plan_execute_batch",
)
FCC_NOALLOC_SCRATCH_BYTES = 64 << 20


class FccTensor(C.Structure):
    _fields_ = [
        ("order", C.c_int32),
        ("nnz", C.c_int64),
        ("coords", C.c_void_p),
        ("vals", C.c_void_p),
        ("extents", C.POINTER(C.c_uint32)),
        ("on_device", C.c_int32),
    ]


class FccOptions(C.Structure):
    _fields_ = [
        ("planner", C.c_int32),
        ("tile_hint", C.c_int32),
        ("sink", C.c_int32),
        ("force_acc", C.c_int32),
        ("col_window", C.c_int32),
        ("workgroups", C.c_int32),
        ("collect_stats", C.c_int32),
        ("left_tile_stride", C.c_int32),
        ("summarize", C.c_int32),
        ("dense_keys", C.c_float),
        ("mirror", C.c_int32),
        ("judge", C.c_int32),
    ]


class FccStats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_index", C.c_double),
        ("ms_plan", C.c_double),
        ("ms_symbolic", C.c_double),
        ("ms_numeric", C.c_double),
        ("ms_overflow", C.c_double),
        ("products", C.c_uint64),
        ("bytes_alg", C.c_uint64),
        ("left_tiles", C.c_uint64),
        ("units", C.c_uint64),
        ("units_dense", C.c_uint64),
        ("units_lhash", C.c_uint64),
        ("units_ghash", C.c_uint64),
        ("ghash_entries", C.c_uint64),
        ("reruns", C.c_uint64),
        ("row_tile", C.c_int32),
        ("col_window", C.c_int32),
        ("mlp_mape", C.c_double),
        ("ms_numeric_kernel", C.c_double),
        ("ms_done", C.c_double),
        ("units_direct", C.c_uint64),
        ("dense_keys", C.c_uint64),
        ("dense_key_products", C.c_uint64),
        ("ms_dense", C.c_double),
        ("ms_dense_gemm", C.c_double),
        ("dense_flops", C.c_uint64),
        ("dense_block_bytes", C.c_uint64),
        ("units_block_counted", C.c_uint64),
        ("mirrored", C.c_int32),
        ("mape_tiles", C.c_uint64),
    ]


class FccResult(C.Structure):
    _fields_ = [
        ("order", C.c_int32),
        ("nnz", C.c_int64),
        ("coords", C.c_void_p),
        ("coord_stride", C.c_int64),
        ("vals", C.c_void_p),
        ("extents", C.c_uint32 * FCC_MAX_ORDER),
        ("coord_hash", C.c_uint64),
        ("val_sum", C.c_double),
        ("val_wsum", C.c_double),
        ("stats", FccStats),
        ("priv", C.c_void_p),
    ]


class FccPlanInfo(C.Structure):
    _fields_ = [
        ("out_nnz", C.c_int64),
        ("products", C.c_uint64),
        ("left_tiles", C.c_uint64),
        ("units", C.c_uint64),
        ("units_dense", C.c_uint64),
        ("units_lhash", C.c_uint64),
        ("units_ghash", C.c_uint64),
        ("row_tile", C.c_int32),
        ("col_window", C.c_int32),
        ("order", C.c_int32),
        ("extents", C.c_uint32 * FCC_MAX_ORDER),
        ("nnz_lower_bound", C.c_uint64),
        ("nnz_upper_bound", C.c_uint64),
        ("ms_plan", C.c_double),
    ]


class FccTuneEntry(C.Structure):
    _fields_ = [
        ("bucket", C.c_int32),
        ("row_tile", C.c_int32),
        ("col_window", C.c_int32),
        ("workgroups", C.c_int32),
        ("ms", C.c_double),
    ]


_lib = None


def load(path: str | None = None):
    """Load libfastcc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    P = C.c_void_p
    lib.fcc_create.argtypes = [C.POINTER(P), C.c_int]
    lib.fcc_create.restype = C.c_int
    lib.fcc_destroy.argtypes = [P]
    lib.fcc_destroy.restype = None
    lib.fcc_last_error.argtypes = [P]
    lib.fcc_last_error.restype = C.c_char_p
    lib.fcc_stream.argtypes = [P]
    lib.fcc_stream.restype = P
    lib.fcc_options_default.argtypes = [C.POINTER(FccOptions)]
    lib.fcc_options_default.restype = None
    lib.fcc_contract.argtypes = [P, C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(FccTensor),
                                 C.POINTER(C.c_int32), C.c_int32, C.POINTER(FccOptions), C.POINTER(FccResult)]
    lib.fcc_contract.restype = C.c_int
    lib.fcc_contract_ex.argtypes = [P, C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                    C.c_int32, C.POINTER(FccOptions), C.POINTER(FccResult)]
    lib.fcc_contract_ex.restype = C.c_int
    lib.fcc_contract_batch.argtypes = [P, C.c_int32, C.POINTER(FccTensor), C.POINTER(C.c_int32),
                                       C.POINTER(FccTensor), C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(FccOptions), C.POINTER(FccResult)]
    lib.fcc_contract_batch.restype = C.c_int
    lib.fcc_result_to_host.argtypes = [P, C.POINTER(FccResult), P, P]
    lib.fcc_result_to_host.restype = C.c_int
    lib.fcc_result_free.argtypes = [P, C.POINTER(FccResult)]
    lib.fcc_result_free.restype = None
    lib.fcc_mlp_param_count.argtypes = []
    lib.fcc_mlp_param_count.restype = C.c_int64
    lib.fcc_mlp_set_params.argtypes = [P, P]
    lib.fcc_mlp_set_params.restype = C.c_int
    lib.fcc_mlp_get_params.argtypes = [P, P]
    lib.fcc_mlp_get_params.restype = C.c_int
    lib.fcc_tile_stats_count.argtypes = [P]
    lib.fcc_tile_stats_count.restype = C.c_int64
    lib.fcc_tile_stats_copy.argtypes = [P, P, P, C.c_int64]
    lib.fcc_tile_stats_copy.restype = C.c_int
    lib.fcc_mlp_train.argtypes = [P, P, P, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                  C.POINTER(C.c_float)]
    lib.fcc_mlp_train.restype = C.c_int
    lib.fcc_mlp_forward_backward.argtypes = [P, P, P, C.c_int64, P, C.POINTER(C.c_float), P]
    lib.fcc_mlp_forward_backward.restype = C.c_int
    lib.fcc_density_bucket.argtypes = [C.c_double]
    lib.fcc_density_bucket.restype = C.c_int32
    lib.fcc_tune_set.argtypes = [P, C.POINTER(FccTuneEntry), C.c_int32]
    lib.fcc_tune_set.restype = C.c_int
    lib.fcc_tune_get.argtypes = [P, C.POINTER(FccTuneEntry), C.c_int32, C.POINTER(C.c_int32)]
    lib.fcc_tune_get.restype = C.c_int
    lib.fcc_result_as_tensor.argtypes = [P, C.POINTER(FccResult), C.POINTER(FccTensor)]
    lib.fcc_result_as_tensor.restype = C.c_int
    lib.fcc_autotune.argtypes = [P, C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(FccTensor),
                                 C.POINTER(C.c_int32), C.c_int32, C.POINTER(FccOptions), C.c_int32,
                                 C.POINTER(FccTuneEntry)]
    lib.fcc_autotune.restype = C.c_int
    lib.fcc_plan_create.argtypes = [P, C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                    C.c_int32, C.POINTER(FccOptions), C.POINTER(P)]
    lib.fcc_plan_create.restype = C.c_int
    lib.fcc_plan_info_get.argtypes = [P, C.POINTER(FccPlanInfo)]
    lib.fcc_plan_info_get.restype = C.c_int
    lib.fcc_plan_tile_count.argtypes = [P]
    lib.fcc_plan_tile_count.restype = C.c_int64
    lib.fcc_plan_tiles.argtypes = [P, P, P, P, C.c_int64]
    lib.fcc_plan_tiles.restype = C.c_int
    lib.fcc_plan_execute.argtypes = [P, P, P, C.c_int32, C.c_int32, C.POINTER(FccResult)]
    lib.fcc_plan_execute.restype = C.c_int
    lib.fcc_plan_create_batch.argtypes = [P, C.c_int32, C.POINTER(FccTensor), C.POINTER(C.c_int32),
                                          C.POINTER(FccTensor), C.POINTER(C.c_int32), C.c_int32,
                                          C.POINTER(FccOptions), C.POINTER(P)]
    lib.fcc_plan_create_batch.restype = C.c_int
    lib.fcc_plan_execute_batch.argtypes = [P, P, P, C.c_int32, C.c_int32, C.POINTER(FccResult)]
    lib.fcc_plan_execute_batch.restype = C.c_int
    lib.fcc_plan_destroy.argtypes = [P]
    lib.fcc_plan_destroy.restype = None
    lib.fcc_contract_noalloc.argtypes = [P, C.POINTER(FccTensor), C.POINTER(C.c_int32), C.POINTER(FccTensor),
                                         C.POINTER(C.c_int32), C.c_int32, C.POINTER(FccOptions), P, P, C.c_int64,
                                         C.POINTER(FccResult)]
    lib.fcc_contract_noalloc.restype = C.c_int
    lib.fcc_set_memory_budget.argtypes = [P, C.c_uint64]
    lib.fcc_set_memory_budget.restype = C.c_int
    lib.fcc_memory_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.fcc_memory_stats.restype = C.c_int
    if path is None:
        _lib = lib
    return lib
