"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), and
bench.py's cpu_baseline leg). ctypes bindings of liboracle.so, the CPU
restatement of KushDB's operator semantics (kush_oracle.cc), plus the optional
reference-runtime pin library _ref/libkushref.so.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkushref.so")

# kor_type / kor_agg_kind / kor_cmp (same numbering as kushgpu.h)
SMALLINT, INT, BIGINT, REAL, DATE, TEXT, BOOLEAN, ENUM = range(8)
SUM, MIN, MAX, AVG, COUNT = range(5)
EQ, NEQ, LT, LEQ, GT, GEQ = range(6)

_NP = {SMALLINT: np.int16, INT: np.int32, BIGINT: np.int64, REAL: np.float64, DATE: np.int32,
       BOOLEAN: np.int8, ENUM: np.int32}


class KorCol(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("data", ctypes.c_void_p), ("nulls", ctypes.c_void_p)]


class KorAgg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("input", KorCol)]


class KorSortKey(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("asc", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("nulls", ctypes.c_void_p), ("str_off", ctypes.c_void_p), ("str_heap", ctypes.c_void_p)]


class KorExprCol(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("data", ctypes.c_void_p), ("nulls", ctypes.c_void_p),
                ("offsets", ctypes.c_void_p), ("heap", ctypes.c_void_p), ("dict", ctypes.c_void_p),
                ("dict_off", ctypes.c_void_p), ("dict_n", ctypes.c_int32)]


class RefSortKey(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("asc", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("nulls", ctypes.c_void_p)]


class KorPred(ctypes.Structure):
    _fields_ = [("col", ctypes.c_int32), ("op", ctypes.c_int32), ("ilit", ctypes.c_int64),
                ("flit", ctypes.c_double)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(ORACLE_SO)
        L = _lib
        L.kor_hash_combine.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        L.kor_value_hash.restype = ctypes.c_int64
        L.kor_value_hash.argtypes = [ctypes.c_int32, ctypes.c_int64]
        L.kor_agg_hash.restype = ctypes.c_int64
        L.kor_agg_hash.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
        L.kor_index_hash.restype = ctypes.c_uint64
        L.kor_index_hash.argtypes = [ctypes.c_uint64]
        L.kor_build_date.restype = ctypes.c_int32
        L.kor_build_date.argtypes = [ctypes.c_int32] * 3
        L.kor_split_date.argtypes = [ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int32)] * 3
        L.kor_scan_select.restype = ctypes.c_int64
        L.kor_scan_select.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.c_int32, ctypes.c_void_p]
        L.kor_hash_join.restype = ctypes.c_int64
        L.kor_hash_join.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_void_p]
        L.kor_group_by.restype = ctypes.c_int64
        L.kor_group_by.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                   ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.kor_aggregate.restype = ctypes.c_int32
        L.kor_aggregate.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]
        L.kor_fast_forward.restype = ctypes.c_int32
        L.kor_fast_forward.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        L.kor_intersect_all.restype = ctypes.c_int64
        L.kor_intersect_all.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]
        L.kor_tuple_dedup.restype = ctypes.c_int64
        L.kor_tuple_dedup.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.kor_memory_index_build.restype = ctypes.c_int64
        L.kor_memory_index_build.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.kor_kdbindex_get.restype = ctypes.c_int64
        L.kor_kdbindex_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_int64]
        L.kor_q_scan_agg.restype = ctypes.c_int64
        L.kor_q_scan_agg.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.kor_q_join_count_sum.restype = ctypes.c_int64
        L.kor_q_join_count_sum.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]
        L.kor_q_join_agg.restype = ctypes.c_int64
        L.kor_q_join_agg.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.kor_q_star.restype = ctypes.c_int64
        L.kor_q_star.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
        L.kor_text_pred.restype = ctypes.c_int32
        L.kor_text_pred.argtypes = [ctypes.c_int32, ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int32]
        L.kor_text_select.restype = ctypes.c_int64
        L.kor_text_select.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.kor_text_minmax.restype = ctypes.c_int64
        L.kor_text_minmax.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int32]
        L.kor_expr_eval.restype = ctypes.c_int32
        L.kor_expr_eval.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.kor_order_by.restype = ctypes.c_int32
        L.kor_order_by.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.kor_dtree_scan_agg.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_void_p]
        L.kor_time_join.restype = ctypes.c_int64
        L.kor_time_join.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int64)]
        L.kor_q_chain.restype = ctypes.c_int64
        L.kor_q_chain.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int32)]
        L.kor_det_sum.restype = ctypes.c_double
        L.kor_det_sum.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    return _lib


def time_join(bkeys, pkeys, pvals):
    """(matches, SUM(value) over matches, build seconds, probe seconds) of the
    reference-style chained join table on the CPU, one thread."""
    b = np.ascontiguousarray(bkeys, dtype=np.int64)
    p = np.ascontiguousarray(pkeys, dtype=np.int64)
    v = np.ascontiguousarray(pvals, dtype=np.int64)
    bs, ps, s = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    m = lib().kor_time_join(len(b), b.ctypes.data, len(p), p.ctypes.data, v.ctypes.data, ctypes.byref(bs),
                            ctypes.byref(ps), ctypes.byref(s))
    return m, s.value, bs.value, ps.value


def chain_query(htabs, p0=b"a", p1=b"%e_a%", vmax=50):
    """The five-table chain (t_i.fk = t_{i-1}.id; t0.name STARTS_WITH p0,
    t1.s LIKE p1, t4.v < vmax): (result tuples, MIN(t0.name) bytes, MIN(t4.v))."""
    sizes = np.array([len(t["id"]) for t in htabs], dtype=np.int64)
    fks = [np.zeros(1, np.int32)] + [np.ascontiguousarray(t["fk"], dtype=np.int32) for t in htabs[1:]]
    fkp = (ctypes.c_void_p * 5)(*[f.ctypes.data for f in fks])
    v4 = np.ascontiguousarray(htabs[4]["v"], dtype=np.int32)
    no, nb = htabs[0]["name_off"], htabs[0]["name_heap"]
    so, sb = htabs[1]["s_off"], htabs[1]["s_heap"]
    row, vmin = ctypes.c_int64(), ctypes.c_int32()
    n = lib().kor_q_chain(sizes.ctypes.data, fkp, v4.ctypes.data, no.ctypes.data, nb.ctypes.data, so.ctypes.data,
                          sb.ctypes.data, p0, len(p0), p1, len(p1), vmax, ctypes.byref(row), ctypes.byref(vmin))
    name = bytes(nb[no[row.value]:no[row.value + 1]]) if n else None
    return n, name, (vmin.value if n else None)


def det_sum(values):
    """The deterministic mode's fixed-order FP64 sum of one group's values
    (input order; NULLs passed as -0.0)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib().kor_det_sum(v.ctypes.data, len(v))


def ref_built():
    """Whether the reference-runtime pin library exists (without loading it)."""
    return os.path.exists(REF_SO)


def ref():
    """The reference-runtime pin library, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
        R = _ref
        R.ref_hash_combine.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_int64]
        R.ref_build_date.restype = ctypes.c_int32
        R.ref_build_date.argtypes = [ctypes.c_int32] * 3
        R.ref_split_date.argtypes = [ctypes.c_int32] + [ctypes.POINTER(ctypes.c_int32)] * 3
        R.ref_fast_forward.restype = ctypes.c_int32
        R.ref_fast_forward.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        R.ref_intersect_all.restype = ctypes.c_int64
        R.ref_intersect_all.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]
        R.ref_tuple_dedup.restype = ctypes.c_int64
        R.ref_tuple_dedup.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        R.ref_hash_join_i64.restype = ctypes.c_int64
        R.ref_hash_join_i64.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                        ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        R.ref_sort_rows.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        R.ref_group_by_i64.restype = ctypes.c_int64
        R.ref_group_by_i64.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return _ref


def _p(a):
    return a.ctypes.data if a is not None else None


class Col:
    """numpy column (+ optional int8 nulls) for the oracle."""

    def __init__(self, typ, data, nulls=None, dictionary=None):
        self.type = typ
        self.data = np.ascontiguousarray(data, dtype=_NP[typ])
        self.nulls = None if nulls is None else np.ascontiguousarray(nulls, dtype=np.int8)
        self.dictionary = dictionary  # ENUM: strings by code

    def kor(self):
        return KorCol(self.type, self.data.ctypes.data, _p(self.nulls))


def hash_combine(h, v):
    x = ctypes.c_int32(h)
    lib().kor_hash_combine(ctypes.byref(x), int(v))
    return x.value


def build_date(y, m, d):
    return lib().kor_build_date(y, m, d)


def split_date(jd):
    y, m, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib().kor_split_date(jd, ctypes.byref(y), ctypes.byref(m), ctypes.byref(d))
    return y.value, m.value, d.value


def scan_select(cols, preds, conjunction=True):
    n = len(cols[0].data)
    ca = (KorCol * len(cols))(*[c.kor() for c in cols])
    pa = (KorPred * max(len(preds), 1))(*[
        KorPred(ci, op, 0 if cols[ci].type == REAL else int(l), float(l) if cols[ci].type == REAL else 0.0)
        for ci, op, l in preds])
    out = np.empty(max(n, 1), dtype=np.int32)
    k = lib().kor_scan_select(n, ca, len(preds), pa, 1 if conjunction else 0, out.ctypes.data)
    return out[:k].copy()


def hash_join(build_keys, probe_keys, build_rows=None, probe_rows=None):
    """build_keys/probe_keys: lists of Col. Returns (probe_rows, build_rows) pairs in reference order."""
    nk = len(build_keys)
    ba = (KorCol * nk)(*[c.kor() for c in build_keys])
    pa = (KorCol * nk)(*[c.kor() for c in probe_keys])
    nb = len(build_rows) if build_rows is not None else len(build_keys[0].data)
    np_ = len(probe_rows) if probe_rows is not None else len(probe_keys[0].data)
    br = None if build_rows is None else np.ascontiguousarray(build_rows, dtype=np.int32)
    pr = None if probe_rows is None else np.ascontiguousarray(probe_rows, dtype=np.int32)
    cap = max(np_, 16)
    while True:
        op = np.empty(cap, dtype=np.int32)
        ob = np.empty(cap, dtype=np.int32)
        k = lib().kor_hash_join(nk, ba, nb, _p(br), pa, np_, _p(pr), cap, op.ctypes.data, ob.ctypes.data)
        if k <= cap:
            return op[:k].copy(), ob[:k].copy()
        cap = k


def group_by(keys, aggs, rows=None, max_groups=1 << 20):
    """keys: list of Col; aggs: list of (kind, Col or None for COUNT(*)).

    Returns (key_vals int64 [G,K], key_nulls, vals int64 [G,A], val_nulls) in insertion order.
    REAL values come back as IEEE bits (view as float64)."""
    n = len(rows) if rows is not None else (len(keys[0].data) if keys else
                                            len(next(c for _, c in aggs if c is not None).data))
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    ka = (KorCol * max(len(keys), 1))(*[c.kor() for c in keys])
    aa = (KorAgg * max(len(aggs), 1))(*[KorAgg(k, c.kor() if c is not None else KorCol(INT, None, None))
                                        for k, c in aggs])
    ok = np.zeros((max_groups, max(len(keys), 1)), dtype=np.int64)
    okn = np.zeros((max_groups, max(len(keys), 1)), dtype=np.int8)
    ov = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int64)
    ovn = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int8)
    g = lib().kor_group_by(n, _p(rr), len(keys), ka, len(aggs), aa, max_groups, ok.ctypes.data, okn.ctypes.data,
                           ov.ctypes.data, ovn.ctypes.data)
    g = min(g, max_groups)
    return ok[:g, :len(keys)], okn[:g, :len(keys)], ov[:g, :len(aggs)], ovn[:g, :len(aggs)]


def aggregate(aggs, n, rows=None):
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    nn = len(rr) if rr is not None else n
    aa = (KorAgg * max(len(aggs), 1))(*[KorAgg(k, c.kor() if c is not None else KorCol(INT, None, None))
                                        for k, c in aggs])
    ov = np.zeros(max(len(aggs), 1), dtype=np.int64)
    ovn = np.zeros(max(len(aggs), 1), dtype=np.int8)
    got = lib().kor_aggregate(nn, _p(rr), len(aggs), aa, ov.ctypes.data, ovn.ctypes.data)
    return got, ov[:len(aggs)], ovn[:len(aggs)]


def fast_forward(data, prev):
    a = np.ascontiguousarray(data, dtype=np.int32)
    return lib().kor_fast_forward(a.ctypes.data if len(a) else None, len(a), prev)


def intersect_all(lists, next_tuple=0, chunk=64, fn=None):
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in lists]
    ptrs = (ctypes.c_void_p * max(len(arrs), 1))(*[a.ctypes.data for a in arrs])
    sizes = (ctypes.c_int32 * max(len(arrs), 1))(*[len(a) for a in arrs])
    cap = max([len(a) for a in arrs] + [1])
    out = np.empty(cap, dtype=np.int32)
    f = fn or lib().kor_intersect_all
    k = f(len(arrs), ptrs, sizes, next_tuple, chunk, out.ctypes.data, cap)
    return out[:k].copy()


def tuple_dedup(tuples, fn=None):
    t = np.ascontiguousarray(tuples, dtype=np.int32)
    n, ln = t.shape
    is_new = np.zeros(max(n, 1), dtype=np.uint8)
    out = np.zeros((max(n, 1), ln), dtype=np.int32)
    f = fn or lib().kor_tuple_dedup
    k = f(n, ln, t.ctypes.data, is_new.ctypes.data, out.ctypes.data)
    return is_new[:n].astype(bool), out[:k].copy()


def kdbindex_get(file_bytes, key_type, key):
    buf = np.frombuffer(file_bytes, dtype=np.uint8)
    if key_type == REAL:
        bits = int(np.array([key], dtype=np.float64).view(np.int64)[0])
    else:
        bits = int(key)
    cap = 1 << 20
    out = np.empty(cap, dtype=np.int32)
    k = lib().kor_kdbindex_get(buf.ctypes.data, key_type, bits, out.ctypes.data, cap)
    return out[:k].copy()


class TextCol:
    """numpy TEXT column: int64 offsets [n+1] + uint8 heap (+ int8 nulls)."""

    def __init__(self, offsets, heap, nulls=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.heap = np.ascontiguousarray(heap if len(heap) else np.zeros(1, np.uint8), dtype=np.uint8)
        self.nulls = None if nulls is None else np.ascontiguousarray(nulls, dtype=np.int8)


def text_select(col, op, pattern, rows=None):
    n = len(rows) if rows is not None else len(col.offsets) - 1
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    out = np.empty(max(n, 1), dtype=np.int32)
    pb = pattern.encode() if isinstance(pattern, str) else bytes(pattern)
    k = lib().kor_text_select(n, col.offsets.ctypes.data, col.heap.ctypes.data, _p(col.nulls), op, pb, len(pb),
                              _p(rr), out.ctypes.data)
    return out[:k].copy()


def text_minmax(col, rows=None, is_max=False):
    n = len(rows) if rows is not None else len(col.offsets) - 1
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    return lib().kor_text_minmax(n, col.offsets.ctypes.data, col.heap.ctypes.data, _p(col.nulls), _p(rr),
                                 1 if is_max else 0)


def _strings(strs):
    enc = [x if isinstance(x, bytes) else x.encode() for x in strs]
    off = np.zeros(len(enc) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(b) for b in enc])
    heap = np.frombuffer(b"".join(enc) + b"\0", dtype=np.uint8).copy()
    return off, heap


def order_by(keys):
    """keys: list of (Col or TextCol-like tuple, asc). A Col of type ENUM
    carries its dictionary as col.dictionary (strings by code); a TEXT key is
    ("text", strings, nulls). Returns row positions in ORDER BY order, or None
    when a key is BOOLEAN (the reference throws)."""
    ks = (KorSortKey * max(len(keys), 1))()
    keep = []
    n = None
    for i, (c, asc) in enumerate(keys):
        if isinstance(c, tuple) and c[0] == "text":
            _, strs, nulls = c
            off, heap = _strings(strs)
            nl = None if nulls is None else np.ascontiguousarray(nulls, dtype=np.int8)
            keep += [off, heap, nl]
            ks[i] = KorSortKey(TEXT, 1 if asc else 0, None, _p(nl), off.ctypes.data, heap.ctypes.data)
            n = len(strs)
        else:
            so = sh = None
            if c.type == ENUM:
                off, heap = _strings(c.dictionary)
                keep += [off, heap]
                so, sh = off.ctypes.data, heap.ctypes.data
            ks[i] = KorSortKey(c.type, 1 if asc else 0, c.data.ctypes.data, _p(c.nulls), so, sh)
            n = len(c.data)
    out = np.empty(max(n, 1), dtype=np.int32)
    if lib().kor_order_by(n, ks, len(keys), out.ctypes.data) != 0:
        return None
    return out[:n].copy()


def ref_sort_rows(keys):
    """The reference's Vector::Sort over row positions (oracle/_ref), for
    typed non-string keys: list of (Col, asc)."""
    ks = (RefSortKey * max(len(keys), 1))()
    for i, (c, asc) in enumerate(keys):
        ks[i] = RefSortKey(c.type, 1 if asc else 0, c.data.ctypes.data, _p(c.nulls))
    n = len(keys[0][0].data)
    out = np.empty(max(n, 1), dtype=np.int32)
    ref().ref_sort_rows(n, ks, len(keys), out.ctypes.data)
    return out[:n].copy()


def expr_eval(nodes_ptr, nn, cols, n, rows=None):
    """Evaluates an expression node array (kor_expr_node layout) over host
    columns: each a Col (ENUM may carry .dictionary) or ("text", strings,
    nulls). Returns (int64 value bits, int8 NULL flags), or None when the
    expression is rejected."""
    cs = (KorExprCol * max(len(cols), 1))()
    keep = []
    for i, c in enumerate(cols):
        if isinstance(c, tuple) and c[0] == "text":
            _, strs, nulls = c
            off, heap = _strings(strs)
            nl = None if nulls is None else np.ascontiguousarray(nulls, dtype=np.int8)
            keep += [off, heap, nl]
            cs[i] = KorExprCol(TEXT, None, _p(nl), off.ctypes.data, heap.ctypes.data, None, None, 0)
        else:
            do = dh = None
            dn = 0
            if getattr(c, "dictionary", None) is not None:
                do, dh = _strings(c.dictionary)
                keep += [do, dh]
                dn = len(c.dictionary)
            cs[i] = KorExprCol(c.type, c.data.ctypes.data, _p(c.nulls), None, None,
                               dh.ctypes.data if dh is not None else None, do.ctypes.data if do is not None else None,
                               dn)
    rr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
    m = len(rr) if rr is not None else n
    out = np.zeros(max(m, 1), dtype=np.int64)
    nulls = np.zeros(max(m, 1), dtype=np.int8)
    if lib().kor_expr_eval(m, cs, len(cols), nodes_ptr, nn, _p(rr), out.ctypes.data, nulls.ctypes.data) != 0:
        return None
    return out[:m], nulls[:m]


def skinner_join(cards, conds, rows=None):
    """Result set of the Skinner multi-way equi-join: every tuple of tuple ids
    (one per table) that satisfies all conditions, each once.

    Restates what the permutable Skinner join emits, independent of the join
    order the UCT agent picks: each table function keeps the tuples of its
    table that match the joined ones on every equality predicate
    (compile/translators/permutable_skinner_join_translator.cc:359-433 GetBucket
    per equality predicate, :663-832 full scan without an index, flagged
    predicates evaluated once all their tables are joined), NULL keys are not
    indexed and match nothing (:972-994), and the valid-tuple handler inserts
    each full tuple into the TupleIdxTable and emits it only when new
    (:842-912). cards[t] = tuples of table t; conds = [(lt, lkeys, lnulls, rt,
    rkeys, rnulls)] with key arrays indexed by row; rows[t] maps tuple id -> row
    (None: identity). Returns a sorted list of int tuples."""
    nt = len(cards)
    rows = rows or [None] * nt

    def key(c, side, t, tid):
        lt, lk, ln, rt, rk, rn = conds[c]
        ks, ns = (lk, ln) if side == 0 else (rk, rn)
        r = tid if rows[t] is None else int(rows[t][tid])
        if ns is not None and ns[r]:
            return None
        v = ks[r]
        if isinstance(v, (str, bytes)):
            return v
        return float(v) if isinstance(v, (float, np.floating)) else int(v)

    if any(c == 0 for c in cards):
        return []
    # Tables in breadth-first order over the condition graph, one component
    # after the other (the order does not change the set).
    order, seen = [], set()
    for s in range(nt):
        if s in seen:
            continue
        seen.add(s)
        q = [s]
        while q:
            t = q.pop(0)
            order.append(t)
            for c in conds:
                for a, b in ((c[0], c[3]), (c[3], c[0])):
                    if a == t and b not in seen:
                        seen.add(b)
                        q.append(b)
    partial = [{order[0]: i} for i in range(cards[order[0]])]
    for t in order[1:]:
        links = [(ci, 0 if c[0] == t else 1) for ci, c in enumerate(conds)
                 if (c[0] == t and c[3] in partial[0]) or (c[3] == t and c[0] in partial[0])] if partial else []
        nxt = []
        if links:
            ci, side = links[0]
            other = conds[ci][3] if side == 0 else conds[ci][0]
            bucket = {}
            for i in range(cards[t]):
                k = key(ci, side, t, i)
                if k is not None:
                    bucket.setdefault(k, []).append(i)
            for p in partial:
                k = key(ci, 1 - side, other, p[other])
                for i in ([] if k is None else bucket.get(k, [])):
                    q = dict(p)
                    q[t] = i
                    nxt.append(q)
        else:
            nxt = [{**p, t: i} for p in partial for i in range(cards[t])]
        for ci, side in links[1:]:
            other = conds[ci][3] if side == 0 else conds[ci][0]
            keep = []
            for q in nxt:
                a, b = key(ci, side, t, q[t]), key(ci, 1 - side, other, q[other])
                if a is not None and b is not None and a == b:
                    keep.append(q)
            nxt = keep
        partial = nxt
        if not partial:
            return []
    return sorted({tuple(p[t] for t in range(nt)) for p in partial})


def text_hash_join(build, build_nulls, probe, probe_nulls):
    """TEXT-key hash join: (probe_rows, build_rows) of every pair of equal,
    non-NULL strings, probe order then build insertion order
    (compile/translators/hash_join_translator.cc:79-139 with CheckEquality over
    String::Equals, runtime/string.cc:42-46; NULL build keys are skipped,
    :86-101)."""
    buckets = {}
    for i, v in enumerate(build):
        if build_nulls is None or not build_nulls[i]:
            buckets.setdefault(v, []).append(i)
    pr, br = [], []
    for j, v in enumerate(probe):
        if probe_nulls is not None and probe_nulls[j]:
            continue
        for i in buckets.get(v, []):
            pr.append(j)
            br.append(i)
    return np.array(pr, dtype=np.int32), np.array(br, dtype=np.int32)
