"""ctypes binding of libreset_amd.so (the C ABI in include/reset_amd.h).

The library must have been built (``python -m reset_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libreset_amd.so")

RS_OK, RS_EINVAL, RS_ENOMEM, RS_EOVERFLOW, RS_EDEVICE, RS_EBUSY = 0, 1, 2, 3, 4, 5
RS_ECANCELED, RS_ETIMEDOUT = 6, 7
REQ_IR, REQ_SSS = 0, 1
NORMAL, HIERARCHICAL = 0, 1
BUILD_NORMAL, BUILD_HIERARCHICAL, BUILD_BOTH = 1, 2, 3
PAIRS_MAX_BINS = 20
RS_PAIRS_STATIC, RS_PAIRS_STEAL = 0, 16  # rs_pairs_opts.schedule: static, or pairs per work-stealing take


class ResetError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


class rs_tables_desc(C.Structure):
    _fields_ = [
        ("num_sets", C.c_int64), ("num_elements", C.c_int64), ("bucket_num", C.c_int64),
        ("max_length", C.c_int32), ("bucket_size", C.c_int32), ("load_factor", C.c_float),
        ("layouts", C.c_int32), ("d_table_normal", C.c_void_p), ("d_table_hierarchical", C.c_void_p),
        ("d_lengths", C.c_void_p), ("d_table_offsets", C.c_void_p), ("d_offsets", C.c_void_p),
        ("d_elements", C.c_void_p), ("conflicts_normal", C.c_int64),
        ("conflicts_hierarchical", C.c_int64),
    ]


class rs_pairs_bin(C.Structure):
    _fields_ = [("load_factor", C.c_float), ("bucket_size", C.c_int32), ("chunk", C.c_int32),
                ("unroll", C.c_int32), ("lanes", C.c_int32)]


class rs_pairs_opts(C.Structure):
    _fields_ = [("num_bins", C.c_int32), ("bins", rs_pairs_bin * PAIRS_MAX_BINS),
                ("waves_per_cu", C.c_int32), ("lds_slots_per_wave", C.c_int32), ("schedule", C.c_int32)]


class rs_completion(C.Structure):
    _fields_ = [("query_id", C.c_int64), ("count", C.c_int32), ("status", C.c_int32),
                ("submit_ms", C.c_double), ("done_ms", C.c_double), ("kind", C.c_int32),
                ("reserved", C.c_int32)]


_p, _i32, _i64, _u32, _u64, _f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); the complete exported C ABI.
SIGNATURES = {
    "rs_version": (C.c_char_p, []),
    "rs_status_string": (C.c_char_p, [_i32]),
    "rs_last_error": (C.c_char_p, []),
    "rs_calculate_length": (_i32, [_i32, _f32, _i32]),
    "rs_calculate_lengths": (_i32, [_p, _i64, _f32, _i32, _p, _p, _p]),
    "rs_tables_build": (_i32, [_p, _i64, _p, _f32, _i32, _i32, _p, _pp]),
    "rs_tables_rebuild_async": (_i32, [_p, _i32, _p]),
    "rs_tables_timing": (_i32, [_p, _i32]),
    "rs_tables_kernel_ms": (_i32, [_p, C.POINTER(C.c_float)]),
    "rs_ir_set_steal_threshold": (_i32, [_i64]),
    "rs_tables_describe": (_i32, [_p, C.POINTER(rs_tables_desc)]),
    "rs_tables_free": (None, [_p]),
    "rs_tables_status": (_i32, [_p, _p]),
    "rs_expand_rows": (_i32, [_p, _i64, _i64, _p, _p]),
    "rs_tables_probe_order": (_i32, [_p, _i32, _p, _p]),
    "rs_tc_plan_create": (_i32, [_p, _p, _pp]),
    "rs_tc_plan_free": (None, [_p]),
    "rs_tc_plan_work": (_i32, [_p, _p, _p]),
    "rs_tc_plan_staged": (_i32, [_p, _p, _p, _p]),
    "rs_tc_count_async": (_i32, [_p, _p, _i32, _p, _p, _p]),
    "rs_tc_count_plain_async": (_i32, [_p, _p, _i32, _p, _p, _p]),
    "rs_tc_count": (_i32, [_p, _i32, _p, _p, _p]),
    "rs_sss_verify_async": (_i32, [_p, _i32, _p, _p, _p, _i64, _i32, _i32, _p, _p, _p]),
    "rs_sss_selfjoin": (_i32, [_p, _i64, _p, _i32, _i32, _f32, _i32, _p, _pp]),
    "rs_sss_result_info": (_i32, [_p, _p, _p, _pp]),
    "rs_sss_result_timings": (_i32, [_p, _p]),
    "rs_sss_result_work": (_i32, [_p, _p, _p]),
    "rs_sss_result_free": (None, [_p]),
    "rs_sss_index_build": (_i32, [_p, _i64, _p, _i32, _i32, _p, _pp]),
    "rs_sss_index_free": (None, [_p]),
    "rs_sss_lookup": (_i32, [_p, _p, _i64, _p, _p, _pp]),
    "rs_ir_query_async": (_i32, [_p, _i32, _p, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "rs_ir_query_noalloc": (_i32, [_p, _i32, _p, _p, _p, _i64, _p, _p, _p, _p, _p]),
    "rs_server_create": (_i32, [_p, _i32, _i64, _i32, _i32, _pp]),
    "rs_server_submit_ir": (_i32, [_p, _i64, _p, _i32]),
    "rs_server_attach_sss": (_i32, [_p, _p]),
    "rs_server_submit_sss": (_i32, [_p, _i64, _p, _i32]),
    "rs_server_poll": (_i32, [_p, _p, _i32, _p]),
    "rs_server_result": (_i32, [_p, _i64, _p, _i64, _p]),
    "rs_server_drain": (_i32, [_p]),
    "rs_server_cancel": (_i32, [_p, _i64]),
    "rs_server_set_queue_timeout": (_i32, [_p, C.c_double]),
    "rs_server_pause": (_i32, [_p, _i32]),
    "rs_server_stats": (_i32, [_p, _p, _p, _p]),
    "rs_server_free": (None, [_p]),
    "rs_pairs_default_opts": (None, [C.POINTER(rs_pairs_opts)]),
    "rs_pairs_intersect_plain": (_i32, [_p, _p, _p, _p, _p, _p, _i64, _f32, _i32, _p, _p, _p]),
    "rs_pairs_plan_create": (_i32, [_i64, C.POINTER(rs_pairs_opts), _pp]),
    "rs_pairs_plan_free": (None, [_p]),
    "rs_pairs_intersect_batched_async": (_i32, [_p, _p, _p, _p, _p, _i64, _p, _p]),
    "rs_pairs_intersect_indexed_async": (_i32, [_p, _p, _p, _p, _p, _i64, _p, _p]),
    "rs_pairs_intersect_cross_async": (_i32, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p]),
    "rs_pairs_graph_capture": (_i32, [_p, _p, _p, _p, _p, _i64, _p, _p]),
    "rs_pairs_graph_launch": (_i32, [_p, _p]),
    "rs_pairs_plan_status": (_i32, [_p, _p]),
    "rs_pairs_plan_timing": (_i32, [_p, _i32]),
    "rs_pairs_plan_kernel_ms": (_i32, [_p, C.POINTER(C.c_float)]),
    "rs_pairs_autotune": (_i32, [_p, _p, _p, _p, _i64, _i32, C.POINTER(rs_pairs_opts), _p]),
    "rs_hyper_cost": (C.c_double, [_i64, C.c_double, _i32, _i32, _i32, _i32]),
    "rs_hyper_solve": (_i32, [_i64, _i32, _i32, _i32, _p, _p, _p]),
    "rs_gen_pair_sets_async": (_i32, [_p, _p, _p, _i64, _u64, _p, _p, _p]),
    "rs_gen_rmat_async": (_i32, [_i32, _i64, _u32, _u32, _u32, _u64, _p, _p]),
    "rs_comp_build": (_i32, [_p, _i64, _p, _i32, _f32, _p, _pp]),
    "rs_comp_info": (_i32, [_p, C.POINTER(_i64), C.POINTER(_i64)]),
    "rs_comp_array_bytes": (_i32, [_p, _i32, C.POINTER(_i64)]),
    "rs_comp_copy_out": (_i32, [_p, _i32, _p, _i64]),
    "rs_comp_tc_count": (_i32, [_p, C.POINTER(_u64), _p]),
    "rs_comp_sss_verify_async": (_i32, [_p, _p, _p, _i64, _i32, _i32, _p, _p, _p]),
    "rs_comp_ir_query_async": (_i32, [_p, _p, _p, _i64, _p, _p, _p, _p]),
    "rs_comp_status": (_i32, [_p, _p]),
    "rs_comp_free": (None, [_p]),
}
COMP_KINDS = {"linear": 0, "cuckoo": 1, "hopscotch": 2, "roaring": 3}

_lib = None


def load():
    """Load the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m reset_amd.build` "
                          "(the HIP path has no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by missing_symbols(); calling it raises
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def missing_symbols():
    """Entry points of the C ABI that the built library does not export."""
    lib = load()
    return [n for n in SIGNATURES if getattr(lib, n, None) is None]


def check(status):
    if status != RS_OK:
        raise ResetError(status, load().rs_last_error().decode())
    return status
