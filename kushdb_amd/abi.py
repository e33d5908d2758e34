"""ctypes declarations of the kushgpu C ABI (include/kushgpu.h).

Loading fails loudly when libkushgpu.so is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkushgpu.so")

# enum kg_status
KG_OK, KG_ERR_INVALID, KG_ERR_HIP, KG_ERR_OOM, KG_ERR_CAPACITY, KG_ERR_UNSUPPORTED = range(6)
# enum kg_type (catalog::TypeId order)
KG_SMALLINT, KG_INT, KG_BIGINT, KG_REAL, KG_DATE, KG_TEXT, KG_BOOLEAN, KG_ENUM = range(8)
# enum kg_cmp
KG_EQ, KG_NEQ, KG_LT, KG_LEQ, KG_GT, KG_GEQ = range(6)
# enum kg_agg_kind
KG_AGG_SUM, KG_AGG_MIN, KG_AGG_MAX, KG_AGG_AVG, KG_AGG_COUNT = range(5)


class KgCol(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("data", ctypes.c_void_p), ("nulls", ctypes.c_void_p)]


class KgPred(ctypes.Structure):
    _fields_ = [("col", ctypes.c_int32), ("op", ctypes.c_int32),
                ("ilit", ctypes.c_int64), ("flit", ctypes.c_double)]


class KgTextCol(ctypes.Structure):
    _fields_ = [("offsets", ctypes.c_void_p), ("bytes", ctypes.c_void_p), ("nulls", ctypes.c_void_p)]


class KgFmtCol(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("dict_n", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("nulls", ctypes.c_void_p), ("dict", ctypes.c_void_p), ("dict_off", ctypes.c_void_p)]


class KgExprNode(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("type", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32),
                ("c", ctypes.c_int32), ("flags", ctypes.c_int32), ("ilit", ctypes.c_int64), ("flit", ctypes.c_double),
                ("str", ctypes.c_void_p), ("str_len", ctypes.c_int32), ("nvalues", ctypes.c_int32),
                ("values", ctypes.c_void_p)]


class KgEnumDict(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("count", ctypes.c_int32)]


# enum kg_expr_op
(KG_E_COL, KG_E_LIT, KG_E_ADD, KG_E_SUB, KG_E_MUL, KG_E_DIV, KG_E_EQ, KG_E_NEQ, KG_E_LT, KG_E_LEQ, KG_E_GT, KG_E_GEQ,
 KG_E_AND, KG_E_OR, KG_E_STARTS_WITH, KG_E_ENDS_WITH, KG_E_CONTAINS, KG_E_NOT, KG_E_IS_NULL, KG_E_CASE,
 KG_E_INT_TO_REAL, KG_E_EXTRACT_YEAR, KG_E_LIKE, KG_E_IN) = range(24)
KG_E_NULL = 1

# enum kg_text_op
(KG_TEXT_EQ, KG_TEXT_NEQ, KG_TEXT_LT, KG_TEXT_LEQ, KG_TEXT_GT, KG_TEXT_GEQ, KG_TEXT_STARTS_WITH, KG_TEXT_ENDS_WITH,
 KG_TEXT_CONTAINS, KG_TEXT_LIKE) = range(10)


class KgAgg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("col", ctypes.c_int32), ("col2", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


# enum kg_query_kind / kg_serve_mode
KG_Q_SCAN_AGG, KG_Q_JOIN_AGG, KG_Q_CHAIN = range(3)
KG_SERVE_PLAIN, KG_SERVE_BATCHED = range(2)


class KgChainTable(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("cols", ctypes.POINTER(KgCol)), ("ncols", ctypes.c_int32),
                ("npreds", ctypes.c_int32), ("preds", ctypes.POINTER(KgPred)),
                ("text", ctypes.POINTER(KgTextCol)), ("text_pattern", ctypes.c_char_p), ("text_op", ctypes.c_int32),
                ("text_plen", ctypes.c_int32), ("id_col", ctypes.c_int32), ("fk_col", ctypes.c_int32),
                ("fk_index", ctypes.c_void_p)]


class KgChain(ctypes.Structure):
    _fields_ = [("ntables", ctypes.c_int32), ("min_text_table", ctypes.c_int32),
                ("tables", ctypes.POINTER(KgChainTable)), ("min_text", ctypes.POINTER(KgTextCol)),
                ("min_int_table", ctypes.c_int32), ("min_int_col", ctypes.c_int32),
                ("skinner_budget", ctypes.c_int64), ("seed", ctypes.c_int64)]


class KgQuery(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("table_id", ctypes.c_int32), ("n", ctypes.c_int64),
                ("cols", ctypes.POINTER(KgCol)), ("ncols", ctypes.c_int32), ("npreds", ctypes.c_int32),
                ("preds", ctypes.POINTER(KgPred)), ("key_cols", ctypes.POINTER(ctypes.c_int32)),
                ("key_card", ctypes.POINTER(ctypes.c_int32)), ("nkeys", ctypes.c_int32), ("naggs", ctypes.c_int32),
                ("aggs", ctypes.POINTER(KgAgg)), ("probe_key_col", ctypes.c_int32),
                ("build_key_col", ctypes.c_int32), ("build_n", ctypes.c_int64),
                ("build_cols", ctypes.POINTER(KgCol)), ("build_ncols", ctypes.c_int32),
                ("build_npreds", ctypes.c_int32), ("build_preds", ctypes.POINTER(KgPred)),
                ("build_code_col", ctypes.c_int32), ("group_card", ctypes.c_int32),
                ("load_factor", ctypes.c_double), ("chain", ctypes.POINTER(KgChain))]


class KgQueryResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("reserved", ctypes.c_int32), ("max_groups", ctypes.c_int64),
                ("ngroups", ctypes.c_int64), ("keys", ctypes.c_void_p), ("vals", ctypes.c_void_p),
                ("nulls", ctypes.c_void_p)]


class KgBatchOpts(ctypes.Structure):
    _fields_ = [("arena_bytes", ctypes.c_int64), ("reserved", ctypes.c_int64)]


class KgServeOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("slots", ctypes.c_int32), ("max_batch", ctypes.c_int32),
                ("lanes", ctypes.c_int32), ("arena_bytes", ctypes.c_int64), ("max_batch_rows", ctypes.c_int64),
                ("lane_scan_rows", ctypes.c_int64)]


class KgServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("qps", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("max_s", ctypes.c_double), ("mean_batch", ctypes.c_double),
                ("batches", ctypes.c_int64), ("fallbacks", ctypes.c_int64), ("lane_scans", ctypes.c_int64)]


class KgJoinTuning(ctypes.Structure):
    _fields_ = [("load_factor", ctypes.c_double), ("probe_wg_per_cu", ctypes.c_int32),
                ("probe_regions", ctypes.c_int32), ("probe_ms", ctypes.c_double), ("tuning_class", ctypes.c_uint64),
                ("from_cache", ctypes.c_int32), ("trials", ctypes.c_int32)]



class KgSkinnerTable(ctypes.Structure):
    _fields_ = [("card", ctypes.c_int64), ("d_rows", ctypes.c_void_p)]


class KgSkinnerCond(ctypes.Structure):
    _fields_ = [("left_table", ctypes.c_int32), ("right_table", ctypes.c_int32), ("left", KgCol), ("right", KgCol)]


class KgSkinnerOpts(ctypes.Structure):
    _fields_ = [("budget_per_episode", ctypes.c_int64), ("seed", ctypes.c_int64), ("max_episodes", ctypes.c_int64),
                ("nprefix", ctypes.c_int32), ("prefix", ctypes.c_int32 * 8)]


class KgSkinnerStats(ctypes.Structure):
    _fields_ = [("episodes", ctypes.c_int64), ("distinct_orders", ctypes.c_int64),
                ("result_tuples", ctypes.c_int64), ("distinct_tuples", ctypes.c_int64),
                ("partial_tuples", ctypes.c_int64), ("last_order", ctypes.c_int32 * 8)]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi8 = ctypes.POINTER(ctypes.c_int8)

# name -> (restype, argtypes); every symbol include/kushgpu.h declares.
SIGNATURES = {
    "kg_ctx_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "kg_ctx_destroy": (_i32, [_vp]),
    "kg_last_error": (ctypes.c_char_p, [_vp]),
    "kg_ctx_set_stream": (_i32, [_vp, _vp]),
    "kg_ctx_synchronize": (_i32, [_vp]),
    "kg_ctx_set_deterministic": (_i32, [_vp, _i32]),
    "kg_ctx_set_build_regions": (_i32, [_vp, _i32]),
    "kg_ctx_set_agg_regions": (_i32, [_vp, _i32]),
    "kg_ctx_last_path": (ctypes.c_char_p, [_vp]),
    "kg_ctx_set_probe_regions": (_i32, [_vp, _i32]),
    "kg_ctx_set_timing": (_i32, [_vp, _i32]),
    "kg_ctx_agg_order": (_i32, [_vp, _i64, _pi32, _pi32]),
    "kg_ctx_kernel_stats": (_i32, [_vp, ctypes.POINTER(ctypes.c_double), _pi64, _i32]),
    "kg_scan_select": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgPred), _i32, _i32,
                              _vp, _pi64]),
    "kg_scan_select_noalloc": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgPred), _i32, _i32,
                                      _vp, _pi64]),
    "kg_scan_group_agg": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgPred), _i32,
                                 _pi32, _pi32, _i32, ctypes.POINTER(KgAgg), _i32, _i64, _pi64, _pi32, _pi64,
                                 _pi8]),
    "kg_ht_create": (_i32, [_vp, _i64, ctypes.c_double, ctypes.POINTER(_vp)]),
    "kg_ht_build": (_i32, [_vp, _vp, ctypes.POINTER(KgCol), _vp, _i64]),
    "kg_ht_probe": (_i32, [_vp, _vp, ctypes.POINTER(KgCol), _vp, _i64, _vp, _vp, _i64, _pi64]),
    "kg_ht_probe_count": (_i32, [_vp, _vp, ctypes.POINTER(KgCol), _vp, _i64, _pi64]),
    "kg_ht_probe_agg": (_i32, [_vp, _vp, ctypes.POINTER(KgCol), _vp, _i64, _i64, _vp, _vp, _vp, _vp, _pi64]),
    "kg_ht_size": (_i32, [_vp, _vp, _pi64]),
    "kg_join_tune": (_i32, [_vp, ctypes.POINTER(KgCol), _i64, ctypes.POINTER(KgCol), _i64,
                            ctypes.POINTER(KgJoinTuning)]),
    "kg_ht_free": (_i32, [_vp, _vp]),
    "kg_scan_join_group_agg": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgPred), _i32, _vp,
                                      _i32, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgAgg), _i32, _i64, _pi64,
                                      _pi32, _pi64, _pi8]),
    "kg_index_build": (_i32, [_vp, ctypes.POINTER(KgCol), _vp, _i64, ctypes.POINTER(_vp)]),
    "kg_index_info": (_i32, [_vp, _vp, ctypes.POINTER(_vp), _pi64, _pi64]),
    "kg_index_lookup": (_i32, [_vp, _vp, ctypes.POINTER(KgCol), _vp, _i64, _vp, _vp]),
    "kg_index_free": (_i32, [_vp, _vp]),
    "kg_file_column": (_i32, [_vp, ctypes.c_char_p, _i32, ctypes.POINTER(KgCol), _pi64]),
    "kg_file_text": (_i32, [_vp, ctypes.c_char_p, ctypes.POINTER(KgTextCol), _pi64]),
    "kg_file_release": (_i32, [_vp, ctypes.c_char_p]),
    "kg_file_enum": (_i32, [_vp, ctypes.c_char_p, ctypes.c_char_p, _i64, _pi64, _i32, _pi32]),
    "kg_file_index": (_i32, [_vp, ctypes.c_char_p, _i32, ctypes.POINTER(_vp)]),
    "kg_order_by": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, _pi32, ctypes.POINTER(_vp),
                           ctypes.POINTER(KgTextCol), _vp]),
    "kg_enum_ranks": (_i32, [_vp, _pi64, _i32, _pi32]),
    "kg_format_rows": (_i32, [ctypes.POINTER(KgFmtCol), _i32, _i64, _vp, _vp, _i64, _pi64]),
    "kg_expr_eval": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), ctypes.POINTER(KgTextCol), ctypes.POINTER(KgEnumDict),
                            _i32, ctypes.POINTER(KgExprNode), _i32, _vp, ctypes.POINTER(KgCol)]),
    "kg_expr_select": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), ctypes.POINTER(KgTextCol),
                              ctypes.POINTER(KgEnumDict), _i32, ctypes.POINTER(KgExprNode), _i32, _vp, _vp, _pi64]),
    "kg_intersect_batch": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _pi64]),
    "kg_tuple_set_create": (_i32, [_vp, _i32, _i64, ctypes.POINTER(_vp)]),
    "kg_skinner_join": (_i32, [_vp, _i32, ctypes.POINTER(KgSkinnerTable), _i32, ctypes.POINTER(KgSkinnerCond),
                               ctypes.POINTER(KgSkinnerOpts), ctypes.POINTER(_vp), ctypes.POINTER(KgSkinnerStats)]),
    "kg_uct_create": (_i32, [_i32, ctypes.POINTER(ctypes.c_uint8), _i64, _pi32, _i32, ctypes.POINTER(_vp)]),
    "kg_uct_next": (_i32, [_vp, _pi32]),
    "kg_uct_reward": (_i32, [_vp, ctypes.c_double]),
    "kg_uct_free": (None, [_vp]),
    "kg_hash_group_agg": (_i32, [_vp, _i64, ctypes.POINTER(KgCol), _i32, ctypes.POINTER(KgPred), _i32, _pi32, _i32,
                                 ctypes.POINTER(KgAgg), _i32, _i64, _pi64, _pi32, _pi64, _pi8]),
    "kg_text_hash": (_i32, [_vp, ctypes.POINTER(KgTextCol), _vp, _i64, _vp, _vp]),
    "kg_text_pairs_equal": (_i32, [_vp, ctypes.POINTER(KgTextCol), _vp, ctypes.POINTER(KgTextCol), _vp, _i64, _vp,
                                   _pi64]),
    "kg_text_hash_verify": (_i32, [_vp, ctypes.POINTER(KgTextCol), _vp, _vp, _i64, _pi64]),
    "kg_chain_run": (_i32, [_vp, ctypes.POINTER(KgChain), _pi64, _pi64, ctypes.POINTER(ctypes.c_int8), _pi64]),
    "kg_tuple_set_insert": (_i32, [_vp, _vp, _vp, _i64, _vp, _pi64]),
    "kg_tuple_set_extract": (_i32, [_vp, _vp, _vp, _i64, _pi64]),
    "kg_tuple_set_free": (_i32, [_vp, _vp]),
    "kg_text_select": (_i32, [_vp, ctypes.POINTER(KgTextCol), _i32, ctypes.c_char_p, _i32, _vp, _i64, _vp, _pi64]),
    "kg_text_minmax": (_i32, [_vp, ctypes.POINTER(KgTextCol), _vp, _i64, _i32, _pi64, ctypes.c_char_p, _i64, _pi64]),
    "kg_index_join_expand": (_i32, [_vp, _i64, _i32, _vp, _i32, ctypes.POINTER(KgCol), _vp, _vp, _i64, _vp, _pi64]),
    "kg_tuple_rows": (_i32, [_vp, _vp, _i32, _i32, _vp, _i64, _vp]),
    "kg_gather": (_i32, [_vp, ctypes.POINTER(KgCol), _vp, _i64, _vp, _vp]),
    "kg_group_by_i64": (_i32, [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _pi64]),
    "kg_batch_create": (_i32, [_vp, ctypes.POINTER(KgBatchOpts), ctypes.POINTER(_vp)]),
    "kg_batch_free": (_i32, [_vp]),
    "kg_batch_footprint": (_i32, [ctypes.POINTER(KgQuery), _pi64]),
    "kg_batch_arena_bytes": (_i64, [_vp]),
    "kg_batch_submit": (_i32, [_vp, ctypes.POINTER(KgQuery), _i32]),
    "kg_batch_poll": (_i32, [_vp, _pi32]),
    "kg_batch_collect": (_i32, [_vp, ctypes.POINTER(KgQueryResult)]),
    "kg_batch_query_ready": (_i32, [_vp, _i32, _pi32]),
    "kg_batch_collect_query": (_i32, [_vp, _i32, ctypes.POINTER(KgQueryResult)]),
    "kg_batch_release": (_i32, [_vp]),
    "kg_batch_run": (_i32, [_vp, ctypes.POINTER(KgQuery), _i32, ctypes.POINTER(KgQueryResult)]),
    "kg_serve": (_i32, [_vp, ctypes.POINTER(KgQuery), _i32, ctypes.POINTER(ctypes.c_double),
                        ctypes.POINTER(KgServeOpts), ctypes.POINTER(KgQueryResult),
                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(KgServeStats)]),
}

# Helpers exported next to the ABI (not part of include/kushgpu.h's contract).
EXTRA_SIGNATURES = {
    "kg_agg_blocks_for": (_i32, [_i64]),
}

_lib = None


def load():
    """Returns the loaded library, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m kushdb_amd.build` "
            "(there is no CPU fallback for the GPU path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in {**SIGNATURES, **EXTRA_SIGNATURES}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class KgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"kushgpu status {status}: {msg}")
        self.status = status
