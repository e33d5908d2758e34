"""Python host layer over the kushgpu C ABI.

torch tensors provide device memory and the stream; every operator runs in
libkushgpu.so (HIP, gfx950). Names follow the reference's operators:
ScanSelect, HashJoin (runtime::HashTable), GroupByAggregate
(runtime::AggregateHashTable).
"""
import ctypes
import os

import numpy as np
import torch

from . import abi
from .abi import KgAgg, KgCol, KgError, KgPred

_TORCH_TYPE = {
    abi.KG_SMALLINT: torch.int16,
    abi.KG_INT: torch.int32,
    abi.KG_BIGINT: torch.int64,
    abi.KG_REAL: torch.float64,
    abi.KG_DATE: torch.int32,
    abi.KG_BOOLEAN: torch.int8,
    abi.KG_ENUM: torch.int32,
}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


_FILE_DTYPES = {abi.KG_SMALLINT: "<i2", abi.KG_INT: "<i4", abi.KG_DATE: "<i4", abi.KG_ENUM: "<i4",
                abi.KG_BIGINT: "<i8", abi.KG_REAL: "<f8", abi.KG_BOOLEAN: "|i1"}


def _device_view(ptr, n, typestr, device):
    """A torch tensor viewing n values at device address ptr (no copy)."""
    if not ptr or n == 0:
        import numpy as _np
        return torch.empty(0, dtype=torch.from_numpy(_np.zeros(0, dtype=typestr)).dtype, device=device)
    return torch.as_tensor(_DeviceArray(ptr, n, typestr), device=device)


class Column:
    """A resident column: typed device tensor + optional int8 NULL tensor."""

    def __init__(self, sql_type, data, nulls=None):
        want = _TORCH_TYPE[sql_type]
        if data.dtype != want:
            raise TypeError(f"column of type {sql_type} needs {want}, got {data.dtype}")
        if not data.is_cuda or not data.is_contiguous():
            raise ValueError("columns must be contiguous device tensors")
        if nulls is not None and (nulls.dtype != torch.int8 or nulls.numel() != data.numel() or not nulls.is_cuda):
            raise ValueError("nulls must be an int8 device tensor of the column's length")
        self.type = sql_type
        self.data = data
        self.nulls = nulls

    def __len__(self):
        return self.data.numel()

    def desc(self):
        return KgCol(self.type, 0, self.data.data_ptr(),
                     self.nulls.data_ptr() if self.nulls is not None else 0)


class Context:
    """One device + stream (kg_ctx). Work is ordered on torch's current stream."""

    def __init__(self, device=0, deterministic=False):
        self.lib = abi.load()
        h = ctypes.c_void_p()
        st = self.lib.kg_ctx_create(device, ctypes.byref(h))
        if st != abi.KG_OK:
            raise KgError(st, "kg_ctx_create failed (is a GPU visible?)")
        self.h = h
        self.device = device
        self.check(self.lib.kg_ctx_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)))
        self.check(self.lib.kg_ctx_set_deterministic(self.h, 1 if deterministic else 0))

    def check(self, st):
        if st != abi.KG_OK:
            raise KgError(st, self.lib.kg_last_error(self.h).decode())
        return st

    def close(self):
        if self.h:
            self.lib.kg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(self.lib.kg_ctx_synchronize(self.h))

    # -- on-disk columns (reference .kdb / .kdbenum / .kdbindex files) --------------
    def file_column(self, path, sql_type):
        """The typed column file at path (and its _null.kdb flags) resident in
        device memory owned by the context (until file_release)."""
        d = abi.KgCol()
        n = ctypes.c_int64()
        self.check(self.lib.kg_file_column(self.h, str(path).encode(), sql_type, ctypes.byref(d), ctypes.byref(n)))
        dev = f"cuda:{self.device}"
        data = _device_view(d.data or 0, n.value, _FILE_DTYPES[sql_type], dev)
        nulls = _device_view(d.nulls, n.value, "|i1", dev) if d.nulls else None
        return Column(sql_type, data, nulls)

    def file_text(self, path):
        """The TEXT column file at path as offsets + bytes (+ NULL flags)."""
        d = abi.KgTextCol()
        n = ctypes.c_int64()
        self.check(self.lib.kg_file_text(self.h, str(path).encode(), ctypes.byref(d), ctypes.byref(n)))
        dev = f"cuda:{self.device}"
        off = _device_view(d.offsets, n.value + 1, "<i8", dev)
        nbytes = int(off[-1].item()) if n.value else 0
        data = _device_view(d.bytes, nbytes, "|u1", dev)
        nulls = _device_view(d.nulls, n.value, "|i1", dev) if d.nulls else None
        return TextColumn(off, data, nulls)

    def file_release(self, path=None):
        self.check(self.lib.kg_file_release(self.h, None if path is None else str(path).encode()))

    def file_enum(self, path):
        """Strings of codes 0..count-1 of a .kdbenum dictionary."""
        cnt = ctypes.c_int32()
        p = str(path).encode()
        self.check(self.lib.kg_file_enum(self.h, p, None, 0, None, 0, ctypes.byref(cnt)))
        cap = max(1, os.path.getsize(path))
        buf = ctypes.create_string_buffer(cap)
        offs = (ctypes.c_int64 * (cnt.value + 1))()
        self.check(self.lib.kg_file_enum(self.h, p, buf, cap, offs, cnt.value, ctypes.byref(cnt)))
        raw = buf.raw
        return [raw[offs[i]:offs[i + 1]].decode() for i in range(cnt.value)]

    def file_index(self, path, key_type):
        """The .kdbindex file at path as a device column index."""
        h = ctypes.c_void_p()
        self.check(self.lib.kg_file_index(self.h, str(path).encode(), key_type, ctypes.byref(h)))
        return ColumnIndex.from_handle(self, h)

    def last_path(self):
        """GROUP BY path of the last aggregation: "register", "lds" or "hashed"."""
        return self.lib.kg_ctx_last_path(self.h).decode()

    def set_build_regions(self, regions):
        """-1: region-partitioned join builds when large; 0: direct inserts; > 0: at least that many regions."""
        self.check(self.lib.kg_ctx_set_build_regions(self.h, regions))

    def set_agg_regions(self, regions):
        """Region count of partitioned probe + GROUP BY (-1 auto, 0 off)."""
        self.check(self.lib.kg_ctx_set_agg_regions(self.h, regions))

    def set_probe_regions(self, regions):
        """Region count of partitioned join probes (-1 auto, 0 direct)."""
        self.check(self.lib.kg_ctx_set_probe_regions(self.h, regions))

    def join_tune(self, build_key, probe_key):
        """Autotuned join settings for this build/probe pair (applied to the
        context); returns a dict with the load factor to create the table with."""
        t = abi.KgJoinTuning()
        bd, pd = build_key.desc(), probe_key.desc()
        self.check(self.lib.kg_join_tune(self.h, ctypes.byref(bd), len(build_key), ctypes.byref(pd), len(probe_key),
                                         ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in abi.KgJoinTuning._fields_}

    def set_timing(self, on):
        self.check(self.lib.kg_ctx_set_timing(self.h, 1 if on else 0))

    def agg_order(self, n):
        """(workgroups, rows per lane step) of the last fixed-order aggregation over n rows."""
        nb = ctypes.c_int32()
        vec = ctypes.c_int32()
        self.check(self.lib.kg_ctx_agg_order(self.h, n, ctypes.byref(nb), ctypes.byref(vec)))
        return nb.value, vec.value

    def kernel_stats(self, reset=True):
        """(summed ms, launches) of the dominant kernels since the last reset."""
        ms = ctypes.c_double()
        k = ctypes.c_int64()
        self.check(self.lib.kg_ctx_kernel_stats(self.h, ctypes.byref(ms), ctypes.byref(k), 1 if reset else 0))
        return ms.value, k.value

    # -- ScanSelect -------------------------------------------------------
    def scan_select(self, cols, preds, conjunction=True, out=None):
        """Row ids (int32, ascending) passing `preds` = [(col_idx, op, literal)]."""
        n = len(cols[0]) if cols else 0
        carr = (KgCol * max(len(cols), 1))(*[c.desc() for c in cols])
        parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
        if out is None:
            out = torch.empty(max(n, 1), dtype=torch.int32, device=f"cuda:{self.device}")
        cnt = ctypes.c_int64()
        self.check(self.lib.kg_scan_select(self.h, n, carr, len(cols), parr, len(preds), 1 if conjunction else 0,
                                           _ptr(out), ctypes.byref(cnt)))
        return out[: cnt.value]

    # -- ScanSelect -> GroupByAggregate over dense code keys ---------------
    def scan_group_agg(self, cols, preds, keys, aggs, max_groups=4096):
        """keys = [(col_idx, cardinality)], aggs = [(kind, col_idx)].

        Returns (codes int32 [G, nkeys], values int64 [G, naggs], nulls int8 [G, naggs])
        in insertion order (first qualifying row of each group)."""
        n = len(cols[0]) if cols else 0
        carr = (KgCol * max(len(cols), 1))(*[c.desc() for c in cols])
        parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
        kc = (ctypes.c_int32 * max(len(keys), 1))(*[k[0] for k in keys])
        kd = (ctypes.c_int32 * max(len(keys), 1))(*[k[1] for k in keys])
        aarr = _aggs(aggs)
        codes = np.zeros((max_groups, max(len(keys), 1)), dtype=np.int32)
        vals = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int64)
        nulls = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int8)
        ng = ctypes.c_int64()
        self.check(self.lib.kg_scan_group_agg(
            self.h, n, carr, len(cols), parr, len(preds), kc, kd, len(keys), aarr, len(aggs), max_groups,
            ctypes.byref(ng), codes.ctypes.data_as(abi._pi32), vals.ctypes.data_as(abi._pi64),
            nulls.ctypes.data_as(abi._pi8)))
        g = ng.value
        return codes[:g, : len(keys)], vals[:g, : len(aggs)], nulls[:g, : len(aggs)]

    # -- GroupByAggregate over hashed keys (any group count) ----------------
    def hash_group_agg(self, cols, preds, key_cols, aggs, max_groups=1 << 16):
        """key_cols = [col_idx] (1..4), aggs = [(kind, col_idx)]. Returns
        (first rows int32 [G], values int64 [G, naggs], nulls int8 [G, naggs])
        in insertion order; a group's keys are the key columns at its first row."""
        n = len(cols[0]) if cols else 0
        carr = (KgCol * max(len(cols), 1))(*[c.desc() for c in cols])
        parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
        kc = (ctypes.c_int32 * max(len(key_cols), 1))(*key_cols)
        aarr = _aggs(aggs)
        first = np.zeros(max(max_groups, 1), dtype=np.int32)
        vals = np.zeros((max(max_groups, 1), max(len(aggs), 1)), dtype=np.int64)
        nulls = np.zeros((max(max_groups, 1), max(len(aggs), 1)), dtype=np.int8)
        ng = ctypes.c_int64()
        self.check(self.lib.kg_hash_group_agg(
            self.h, n, carr, len(cols), parr, len(preds), kc, len(key_cols), aarr, len(aggs), max_groups,
            ctypes.byref(ng), first.ctypes.data_as(abi._pi32), vals.ctypes.data_as(abi._pi64),
            nulls.ctypes.data_as(abi._pi8)))
        g = ng.value
        return first[:g], vals[:g, : len(aggs)], nulls[:g, : len(aggs)]

    # -- GroupByAggregate on a BIGINT key (high cardinality) ----------------
    def group_by_i64(self, keys, vals, max_groups=None, with_first=False):
        n = keys.numel()
        if max_groups is None:
            max_groups = max(n, 1)
        dev = keys.device
        ok = torch.empty(max_groups, dtype=torch.int64, device=dev)
        osum = torch.empty(max_groups, dtype=torch.int64, device=dev)
        ocnt = torch.empty(max_groups, dtype=torch.int64, device=dev)
        ofirst = torch.empty(max_groups, dtype=torch.int32, device=dev) if with_first else None
        ng = ctypes.c_int64()
        self.check(self.lib.kg_group_by_i64(self.h, _ptr(keys), _ptr(vals), n, max_groups, _ptr(ok), _ptr(osum),
                                            _ptr(ocnt), _ptr(ofirst), ctypes.byref(ng)))
        g = ng.value
        return ok[:g], osum[:g], ocnt[:g], (ofirst[:g] if with_first else None)


def _aggs(aggs):
    """[(kind, col)] or [(kind, col, col2)] -> kg_agg array."""
    return (KgAgg * max(len(aggs), 1))(*[KgAgg(a[0], a[1], a[2] if len(a) > 2 else -1, 0) for a in aggs])


def _pred(cols, p):
    ci, op, lit = p
    if cols[ci].type == abi.KG_REAL:
        return KgPred(ci, op, 0, float(lit))
    return KgPred(ci, op, int(lit), 0.0)


class HashTable:
    """HBM open-addressing join table (replaces runtime::HashTable)."""

    def __init__(self, ctx, capacity_rows, load_factor=0.5):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.kg_ht_create(ctx.h, capacity_rows, load_factor, ctypes.byref(h)))
        self.h = h

    def insert(self, key_col, rows=None, nrows=None):
        """Build side: insert key_col's rows (or the selected `rows`)."""
        if nrows is None:
            nrows = rows.numel() if rows is not None else len(key_col)
        d = key_col.desc()
        self.ctx.check(self.ctx.lib.kg_ht_build(self.ctx.h, self.h, ctypes.byref(d), _ptr(rows), nrows))

    def probe(self, key_col, rows=None, out_cap=None, out=None):
        """All (probe row, build row) pairs with equal keys (unordered); out =
        (probe rows, build rows) int32 device tensors of out_cap entries to
        reuse, or None to allocate them."""
        nrows = rows.numel() if rows is not None else len(key_col)
        if out_cap is None:
            out_cap = self.probe_count(key_col, rows)
        dev = key_col.data.device
        if out is not None:
            op, ob = out
            if op.numel() < out_cap or ob.numel() < out_cap:
                raise ValueError("output tensors smaller than out_cap")
        else:
            op = torch.empty(max(out_cap, 1), dtype=torch.int32, device=dev)
            ob = torch.empty(max(out_cap, 1), dtype=torch.int32, device=dev)
        cnt = ctypes.c_int64()
        d = key_col.desc()
        self.ctx.check(self.ctx.lib.kg_ht_probe(self.ctx.h, self.h, ctypes.byref(d), _ptr(rows), nrows, _ptr(op),
                                                _ptr(ob), out_cap, ctypes.byref(cnt)))
        return op[: cnt.value], ob[: cnt.value]

    def probe_count(self, key_col, rows=None):
        nrows = rows.numel() if rows is not None else len(key_col)
        cnt = ctypes.c_int64()
        d = key_col.desc()
        self.ctx.check(self.ctx.lib.kg_ht_probe_count(self.ctx.h, self.h, ctypes.byref(d), _ptr(rows), nrows,
                                                      ctypes.byref(cnt)))
        return cnt.value

    def probe_agg(self, key_col, vals, max_groups, with_first=False):
        """GROUP BY matched build key: (key, SUM(vals), COUNT(*)[, first probe row])."""
        dev = key_col.data.device
        ok = torch.empty(max(max_groups, 1), dtype=torch.int64, device=dev)
        osum = torch.empty(max(max_groups, 1), dtype=torch.int64, device=dev)
        ocnt = torch.empty(max(max_groups, 1), dtype=torch.int64, device=dev)
        ofirst = torch.empty(max(max_groups, 1), dtype=torch.int32, device=dev) if with_first else None
        ng = ctypes.c_int64()
        d = key_col.desc()
        self.ctx.check(self.ctx.lib.kg_ht_probe_agg(self.ctx.h, self.h, ctypes.byref(d), _ptr(vals), len(key_col),
                                                    max_groups, _ptr(ok), _ptr(osum), _ptr(ocnt), _ptr(ofirst),
                                                    ctypes.byref(ng)))
        g = ng.value
        return ok[:g], osum[:g], ocnt[:g], (ofirst[:g] if with_first else None)

    def scan_probe_group_agg(self, cols, preds, probe_key_col, build_codes, group_card, aggs, max_groups=4096):
        """ScanSelect(cols, preds) -> probe cols[probe_key_col] -> GROUP BY
        build_codes[build row] with aggs over probe-side columns. Returns
        (codes [G], values int64 [G, naggs], nulls [G, naggs]) in insertion order."""
        ctx = self.ctx
        n = len(cols[0]) if cols else 0
        carr = (KgCol * max(len(cols), 1))(*[c.desc() for c in cols])
        parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
        codes = np.zeros(max_groups, dtype=np.int32)
        vals = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int64)
        nulls = np.zeros((max_groups, max(len(aggs), 1)), dtype=np.int8)
        ng = ctypes.c_int64()
        bc = ctypes.byref(build_codes.desc()) if build_codes is not None else None
        ctx.check(ctx.lib.kg_scan_join_group_agg(
            ctx.h, n, carr, len(cols), parr, len(preds), self.h, probe_key_col, bc, group_card,
            _aggs(aggs), len(aggs), max_groups, ctypes.byref(ng), codes.ctypes.data_as(abi._pi32),
            vals.ctypes.data_as(abi._pi64), nulls.ctypes.data_as(abi._pi8)))
        g = ng.value
        return codes[:g], vals[:g, : len(aggs)], nulls[:g, : len(aggs)]

    def size(self):
        v = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.kg_ht_size(self.ctx.h, self.h, ctypes.byref(v)))
        return v.value

    def free(self):
        if self.h:
            self.ctx.lib.kg_ht_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ColumnIndex:
    """Device key -> ascending tuple-id buckets (replaces runtime::MemoryColumnIndex)."""

    def __init__(self, ctx, key_col, rows=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        nrows = rows.numel() if rows is not None else len(key_col)
        d = key_col.desc()
        ctx.check(ctx.lib.kg_index_build(ctx.h, ctypes.byref(d), _ptr(rows), nrows, ctypes.byref(h)))
        self.h = h
        p = ctypes.c_void_p()
        nt = ctypes.c_int64()
        nk = ctypes.c_int64()
        ctx.check(ctx.lib.kg_index_info(ctx.h, self.h, ctypes.byref(p), ctypes.byref(nt), ctypes.byref(nk)))
        self.tids_ptr = p.value or 0
        self.ntids = nt.value
        self.nkeys = nk.value

    @classmethod
    def from_handle(cls, ctx, h):
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = h
        p = ctypes.c_void_p()
        nt = ctypes.c_int64()
        nk = ctypes.c_int64()
        ctx.check(ctx.lib.kg_index_info(ctx.h, self.h, ctypes.byref(p), ctypes.byref(nt), ctypes.byref(nk)))
        self.tids_ptr = p.value or 0
        self.ntids = nt.value
        self.nkeys = nk.value
        return self

    def lookup(self, key_col, rows=None):
        """Per probe key: (device pointers int64 tensor, sizes int32 tensor)."""
        n = rows.numel() if rows is not None else len(key_col)
        dev = key_col.data.device
        ptrs = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        cnts = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        d = key_col.desc()
        self.ctx.check(self.ctx.lib.kg_index_lookup(self.ctx.h, self.h, ctypes.byref(d), _ptr(rows), n, _ptr(ptrs),
                                                    _ptr(cnts)))
        return ptrs[:n], cnts[:n]

    def offsets_of(self, ptrs):
        """Device pointers from lookup() as element offsets into the tid array."""
        return (ptrs - self.tids_ptr) // 4

    def tids(self):
        """The tid array as a tensor view (device)."""
        if self.ntids == 0:
            return torch.empty(0, dtype=torch.int32, device=f"cuda:{self.ctx.device}")
        return _device_view_i32(self.tids_ptr, self.ntids, self.ctx.device)

    def free(self):
        if self.h:
            self.ctx.lib.kg_index_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DeviceArray:
    """__cuda_array_interface__ over library-owned device memory (no copy)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 2}


def _device_view_i32(ptr, n, device):
    """int32 tensor view of n values at a library-owned device address."""
    return torch.as_tensor(_DeviceArray(ptr, n, "<i4"), device=f"cuda:{device}")


def intersect_batch(ctx, ptrs, sizes, k, next_tuple=None, out_cap=None):
    """ptrs/sizes: [ntasks*k] device int64 pointers / int32 sizes. Returns
    (offsets [ntasks+1] int64, ids int32)."""
    ntasks = sizes.numel() // k
    dev = sizes.device
    offs = torch.empty(ntasks + 1, dtype=torch.int64, device=dev)
    if out_cap is None:
        out_cap = int(sizes.view(ntasks, k).min(dim=1).values.sum().item()) if ntasks else 0
    out = torch.empty(max(out_cap, 1), dtype=torch.int32, device=dev)
    tot = ctypes.c_int64()
    ctx.check(ctx.lib.kg_intersect_batch(ctx.h, ntasks, k, _ptr(ptrs), _ptr(sizes), _ptr(next_tuple), _ptr(offs),
                                         _ptr(out), out_cap, ctypes.byref(tot)))
    return offs, out[: tot.value]


class TupleSet:
    """Device set of int32[len] tuples (replaces runtime::TupleIdxTable)."""

    def __init__(self, ctx, length, capacity):
        self.ctx = ctx
        self.len = length
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.kg_tuple_set_create(ctx.h, length, capacity, ctypes.byref(h)))
        self.h = h
        self.size = 0

    def insert(self, tuples):
        """tuples: [n, len] int32 device tensor. Returns is_new (uint8 tensor)."""
        n = tuples.shape[0]
        is_new = torch.empty(max(n, 1), dtype=torch.uint8, device=tuples.device)
        k = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.kg_tuple_set_insert(self.ctx.h, self.h, _ptr(tuples.contiguous()), n,
                                                        _ptr(is_new), ctypes.byref(k)))
        self.size += k.value
        return is_new[:n]

    def extract(self):
        out = torch.empty((max(self.size, 1), self.len), dtype=torch.int32, device=f"cuda:{self.ctx.device}")
        k = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.kg_tuple_set_extract(self.ctx.h, self.h, _ptr(out), max(self.size, 1),
                                                         ctypes.byref(k)))
        return out[: k.value]

    def free(self):
        if self.h:
            self.ctx.lib.kg_tuple_set_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


    @classmethod
    def from_handle(cls, ctx, h, length, size):
        """Wraps a set the library created (e.g. by kg_skinner_join)."""
        s = cls.__new__(cls)
        s.ctx, s.len, s.h, s.size = ctx, length, h, size
        return s


def skinner_join(ctx, tables, conds, budget=0, seed=1, prefix=(), max_episodes=0):
    """Adaptive multi-way equi-join (kg_skinner_join). tables: list of
    (card, rows) where rows is an int32 device tensor mapping child tuple ->
    key-column row, or None for the identity; conds: list of
    (left_table, left Column, right_table, right Column). Returns
    (distinct result tuples [n, ntables] int32 device tensor, stats dict)."""
    nt = len(tables)
    tab = (abi.KgSkinnerTable * nt)(*[abi.KgSkinnerTable(int(c), r.data_ptr() if r is not None else 0)
                                      for c, r in tables])
    cs = (abi.KgSkinnerCond * max(len(conds), 1))(*[abi.KgSkinnerCond(lt, rt, lc.desc(), rc.desc())
                                                    for lt, lc, rt, rc in conds])
    o = abi.KgSkinnerOpts(budget, seed, max_episodes, len(prefix), (ctypes.c_int32 * 8)(*prefix))
    st = abi.KgSkinnerStats()
    h = ctypes.c_void_p()
    ctx.check(ctx.lib.kg_skinner_join(ctx.h, nt, tab, len(conds), cs, ctypes.byref(o), ctypes.byref(h),
                                      ctypes.byref(st)))
    ts = TupleSet.from_handle(ctx, h, nt, st.distinct_tuples)
    tuples = ts.extract()
    ts.free()
    stats = {k: getattr(st, k) for k, _ in abi.KgSkinnerStats._fields_ if k != "last_order"}
    stats["last_order"] = list(st.last_order)[:nt]
    return tuples, stats


class UctAgent:
    """The reference's UCT join-order agent (kg_uct_*; host only): next()
    returns the order of the next episode, reward(r) reports its reward."""

    def __init__(self, ntables, edges, seed=1, prefix=()):
        lib = abi.load()
        adj = (ctypes.c_uint8 * (ntables * ntables))()
        for a, b in edges:
            adj[a * ntables + b] = 1
        self.lib, self.n = lib, ntables
        h = ctypes.c_void_p()
        pre = (ctypes.c_int32 * max(len(prefix), 1))(*prefix)
        st = lib.kg_uct_create(ntables, adj, seed, pre, len(prefix), ctypes.byref(h))
        if st != abi.KG_OK:
            raise abi.KgError(st, "kg_uct_create")
        self.h = h

    def next(self):
        order = (ctypes.c_int32 * self.n)()
        st = self.lib.kg_uct_next(self.h, order)
        if st != abi.KG_OK:
            raise abi.KgError(st, "kg_uct_next")
        return list(order)

    def reward(self, r):
        st = self.lib.kg_uct_reward(self.h, r)
        if st != abi.KG_OK:
            raise abi.KgError(st, "kg_uct_reward")

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.kg_uct_free(self.h)
            self.h = None

class TextColumn:
    """A resident TEXT column: int64 offsets [n+1] + uint8 heap (+ int8 NULLs)."""

    def __init__(self, offsets, data, nulls=None):
        if offsets.dtype != torch.int64 or data.dtype != torch.uint8:
            raise TypeError("TEXT columns need int64 offsets and a uint8 heap")
        self.offsets = offsets
        self.data = data if data.numel() else torch.zeros(1, dtype=torch.uint8, device=offsets.device)
        self.nulls = nulls

    @staticmethod
    def from_strings(strings, device="cuda:0", nulls=None):
        """Strings (str, UTF-8 encoded, or raw bytes) as a device column."""
        import numpy as _np
        enc = [s if isinstance(s, (bytes, bytearray)) else s.encode() for s in strings]
        off = _np.zeros(len(enc) + 1, dtype=_np.int64)
        off[1:] = _np.cumsum([len(b) for b in enc])
        heap = _np.frombuffer(b"".join(enc), dtype=_np.uint8) if off[-1] else _np.zeros(1, dtype=_np.uint8)
        return TextColumn(torch.as_tensor(off).to(device), torch.as_tensor(heap.copy()).to(device),
                          None if nulls is None else torch.as_tensor(nulls).to(device))

    def __len__(self):
        return self.offsets.numel() - 1

    def desc(self):
        return abi.KgTextCol(self.offsets.data_ptr(), self.data.data_ptr(),
                             self.nulls.data_ptr() if self.nulls is not None else 0)


def text_select(ctx, col, op, pattern, rows=None):
    """Row ids (input order) whose value satisfies `col op pattern`."""
    n = rows.numel() if rows is not None else len(col)
    out = torch.empty(max(n, 1), dtype=torch.int32, device=col.offsets.device)
    pb = pattern.encode() if isinstance(pattern, str) else bytes(pattern)
    cnt = ctypes.c_int64()
    d = col.desc()
    ctx.check(ctx.lib.kg_text_select(ctx.h, ctypes.byref(d), op, pb, len(pb), _ptr(rows), n, _ptr(out),
                                     ctypes.byref(cnt)))
    return out[: cnt.value]



def text_hash(ctx, col, rows=None):
    """64-bit hashes (int64) and NULL flags (int8) of TEXT values: the keys a
    TEXT equi-join builds and probes (kg_text_hash)."""
    n = rows.numel() if rows is not None else len(col)
    dev = col.offsets.device
    h = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    nl = torch.empty(max(n, 1), dtype=torch.int8, device=dev)
    d = col.desc()
    ctx.check(ctx.lib.kg_text_hash(ctx.h, ctypes.byref(d), _ptr(rows), n, _ptr(h), _ptr(nl)))
    return h[:n], nl[:n]


def text_pairs_equal(ctx, a, rows_a, b, rows_b):
    """Positions i (ascending, int32) where a[rows_a[i]] == b[rows_b[i]] and
    neither is NULL (kg_text_pairs_equal)."""
    n = rows_a.numel()
    out = torch.empty(max(n, 1), dtype=torch.int32, device=a.offsets.device)
    cnt = ctypes.c_int64()
    da, db = a.desc(), b.desc()
    ctx.check(ctx.lib.kg_text_pairs_equal(ctx.h, ctypes.byref(da), _ptr(rows_a.contiguous()), ctypes.byref(db),
                                          _ptr(rows_b.contiguous()), n, _ptr(out), ctypes.byref(cnt)))
    return out[: cnt.value]

def text_hash_verify(ctx, col, hashes, rows=None):
    """Number of values (rows, or all) whose string differs from the string of
    the smallest row with the same hash code (kg_text_hash_verify); 0 means a
    GROUP BY or join on the codes is exact."""
    n = hashes.numel()
    out = ctypes.c_int64()
    d = col.desc()
    ctx.check(ctx.lib.kg_text_hash_verify(ctx.h, ctypes.byref(d), _ptr(rows), _ptr(hashes.contiguous()), n,
                                          ctypes.byref(out)))
    return out.value


def text_minmax(ctx, col, rows=None, is_max=False, cap=4096):
    """(position, bytes) of the first MIN (MAX) non-NULL value, or (-1, None)."""
    n = rows.numel() if rows is not None else len(col)
    pos = ctypes.c_int64()
    ln = ctypes.c_int64()
    buf = ctypes.create_string_buffer(cap)
    d = col.desc()
    ctx.check(ctx.lib.kg_text_minmax(ctx.h, ctypes.byref(d), _ptr(rows), n, 1 if is_max else 0, ctypes.byref(pos),
                                     buf, cap, ctypes.byref(ln)))
    if pos.value < 0:
        return -1, None
    return pos.value, buf.raw[: min(ln.value, cap)]


def index_join_expand(ctx, tuples, src_pos, src_key, src_rows, index, out_cap=None, n=None):
    """Extend [n, width] partial tuples by the index bucket of each tuple's key.
    tuples None: the n one-table tuples (0), ..., (n-1)."""
    if tuples is None:
        width = 1
        dev = f"cuda:{ctx.device}"
    else:
        n, width = tuples.shape
        dev = tuples.device
    total = ctypes.c_int64()
    d = src_key.desc()
    if out_cap is None:
        # First call sizes the output (KG_ERR_CAPACITY reports the total).
        st = ctx.lib.kg_index_join_expand(ctx.h, n, width, _ptr(tuples), src_pos, ctypes.byref(d), _ptr(src_rows),
                                          index.h, 0, ctypes.c_void_p(0), ctypes.byref(total))
        if st not in (abi.KG_OK, abi.KG_ERR_CAPACITY):
            ctx.check(st)
        out_cap = total.value
    out = torch.empty((max(out_cap, 1), width + 1), dtype=torch.int32, device=dev)
    ctx.check(ctx.lib.kg_index_join_expand(ctx.h, n, width, _ptr(tuples), src_pos, ctypes.byref(d), _ptr(src_rows),
                                           index.h, out_cap, _ptr(out), ctypes.byref(total)))
    return out[: total.value]


def tuple_rows(ctx, tuples, pos, sel=None):
    """Row ids of tuple position `pos` (through the table's selection)."""
    n, width = tuples.shape
    out = torch.empty(max(n, 1), dtype=torch.int32, device=tuples.device)
    ctx.check(ctx.lib.kg_tuple_rows(ctx.h, _ptr(tuples), width, pos, _ptr(sel), n, _ptr(out)))
    return out[:n]


def gather(ctx, col, rows):
    """Column `col` materialized at `rows` (a new Column)."""
    n = rows.numel()
    out = torch.empty(max(n, 1), dtype=col.data.dtype, device=col.data.device)
    nl = torch.empty(max(n, 1), dtype=torch.int8, device=col.data.device) if col.nulls is not None else None
    d = col.desc()
    ctx.check(ctx.lib.kg_gather(ctx.h, ctypes.byref(d), _ptr(rows), n, _ptr(out), _ptr(nl)))
    return Column(col.type, out[:n], nl[:n] if nl is not None else None)


# -- batched execution and serving -----------------------------------------------

class Query:
    """One batched pipeline (kg_query). Build with Query.scan_agg / Query.join_agg."""

    def __init__(self, kind, cols, preds, aggs, keys=(), table_id=0, probe_key=-1, build_cols=(), build_key=-1,
                 build_preds=(), build_codes=-1, group_card=1, load_factor=0.0):
        self.kind = kind
        self.cols = list(cols)
        self.build_cols = list(build_cols)
        self.naggs = len(aggs)
        if kind == abi.KG_Q_SCAN_AGG:
            self.nkeys_out = len(keys)
        else:
            self.nkeys_out = 1 if build_codes >= 0 else 0
        carr = (KgCol * max(len(cols), 1))(*[c.desc() for c in cols])
        parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
        kc = (ctypes.c_int32 * max(len(keys), 1))(*[k[0] for k in keys])
        kd = (ctypes.c_int32 * max(len(keys), 1))(*[k[1] for k in keys])
        aarr = _aggs(aggs)
        barr = (KgCol * max(len(build_cols), 1))(*[c.desc() for c in build_cols])
        bparr = (KgPred * max(len(build_preds), 1))(*[_pred(build_cols, p) for p in build_preds])
        self._keep = (carr, parr, kc, kd, aarr, barr, bparr)
        n = len(cols[0]) if cols else 0
        bn = len(build_cols[0]) if build_cols else 0
        self.q = abi.KgQuery(kind, table_id, n, carr, len(cols), len(preds), parr, kc, kd, len(keys), len(aggs),
                             aarr, probe_key, build_key, bn, barr, len(build_cols), len(build_preds), bparr,
                             build_codes, group_card, load_factor)

    @staticmethod
    def scan_agg(cols, preds, keys, aggs, table_id=0):
        """ScanSelect -> GroupByAggregate on dense keys [(col, card)] (as Context.scan_group_agg)."""
        return Query(abi.KG_Q_SCAN_AGG, cols, preds, aggs, keys=keys, table_id=table_id)

    @staticmethod
    def join_agg(cols, preds, probe_key, build_cols, build_key, aggs, build_preds=(), build_codes=-1, group_card=1,
                 table_id=0, load_factor=0.0):
        """ScanSelect(build) -> build; ScanSelect(probe) -> probe -> GROUP BY
        build_cols[build_codes] (or one group when build_codes < 0)."""
        return Query(abi.KG_Q_JOIN_AGG, cols, preds, aggs, table_id=table_id, probe_key=probe_key,
                     build_cols=build_cols, build_key=build_key, build_preds=build_preds, build_codes=build_codes,
                     group_card=group_card if build_codes >= 0 else 1, load_factor=load_factor)

    @staticmethod
    def chain(tables, min_text=None, min_int=None, skinner_budget=0, seed=1, table_id=0):
        """A KG_Q_CHAIN query: tables t0..t(k-1) joined t_i.fk = t_(i-1).id.
        tables: [{"cols": [Column], "preds": [(col, op, lit)], "text": TextColumn
        or None, "text_op": op, "pattern": bytes, "id_col": i, "fk_col": j,
        "fk_index": ColumnIndex over fk_col (all rows) or None}];
        min_text = (table, TextColumn) and min_int = (table, col) name the MIN
        aggregates. Its result is one group: (COUNT of distinct tuples, MIN
        int, row of the MIN string)."""
        self = Query.__new__(Query)
        self.kind = abi.KG_Q_CHAIN
        self.cols, self.build_cols = [], []
        self.naggs, self.nkeys_out = 3, 0
        keep = []
        tarr = (abi.KgChainTable * len(tables))()
        for i, t in enumerate(tables):
            cols = list(t["cols"])
            carr = (KgCol * len(cols))(*[c.desc() for c in cols])
            preds = list(t.get("preds", ()))
            parr = (KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
            keep += [carr, parr]
            tc = t.get("text")
            tdesc = ctypes.pointer(tc.desc()) if tc is not None else None
            pat = bytes(t.get("pattern", b""))
            keep += [tdesc, pat]
            fki = t.get("fk_index")
            keep.append(fki)
            tarr[i] = abi.KgChainTable(len(cols[0]), carr, len(cols), len(preds), parr, tdesc, pat,
                                       t.get("text_op", 0), len(pat), t.get("id_col", -1), t.get("fk_col", -1),
                                       fki.h.value if fki is not None else None)
        mt = ctypes.pointer(min_text[1].desc()) if min_text else None
        keep += [tarr, mt]
        self.chain_desc = abi.KgChain(len(tables), min_text[0] if min_text else -1, tarr, mt,
                                      min_int[0] if min_int else -1, min_int[1] if min_int else -1,
                                      skinner_budget, seed)
        self._keep = tuple(keep)
        self.q = abi.KgQuery()
        self.q.kind = abi.KG_Q_CHAIN
        self.q.table_id = table_id
        self.q.naggs = 3
        self.q.build_code_col = -1
        self.q.probe_key_col = -1
        self.q.build_key_col = -1
        self.q.chain = ctypes.pointer(self.chain_desc)
        return self

    def footprint(self):
        b = ctypes.c_int64()
        abi.load().kg_batch_footprint(ctypes.byref(self.q), ctypes.byref(b))
        return b.value


def chain_run(ctx, query):
    """(COUNT of distinct tuples, MIN int or None, row of the MIN string or
    None) of a Query.chain query (kg_chain_run)."""
    c, v, t = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    nl = ctypes.c_int8()
    ctx.check(ctx.lib.kg_chain_run(ctx.h, ctypes.byref(query.chain_desc), ctypes.byref(c), ctypes.byref(v),
                                   ctypes.byref(nl), ctypes.byref(t)))
    return c.value, (None if nl.value else v.value), (None if t.value < 0 else t.value)


def _result_buffers(queries, max_groups):
    res = (abi.KgQueryResult * max(len(queries), 1))()
    bufs = []
    for i, q in enumerate(queries):
        keys = np.zeros((max_groups, max(q.nkeys_out, 1)), dtype=np.int32)
        vals = np.zeros((max_groups, max(q.naggs, 1)), dtype=np.int64)
        nulls = np.zeros((max_groups, max(q.naggs, 1)), dtype=np.int8)
        res[i] = abi.KgQueryResult(0, 0, max_groups, 0, keys.ctypes.data, vals.ctypes.data, nulls.ctypes.data)
        bufs.append((keys, vals, nulls))
    return res, bufs


def _results(queries, res, bufs):
    out = []
    for i, q in enumerate(queries):
        r = res[i]
        keys, vals, nulls = bufs[i]
        g = r.ngroups if r.status == abi.KG_OK else 0
        out.append((r.status, keys[:g, : q.nkeys_out], vals[:g, : q.naggs], nulls[:g, : q.naggs]))
    return out


class Batch:
    """A batch slot (kg_batch): its own stream, device arena and captured graph."""

    def __init__(self, ctx, arena_bytes=0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        opts = abi.KgBatchOpts(arena_bytes, 0)
        ctx.check(ctx.lib.kg_batch_create(ctx.h, ctypes.byref(opts), ctypes.byref(h)))
        self.h = h
        self._inflight = None

    def submit(self, queries, max_groups=4096):
        qarr = (abi.KgQuery * max(len(queries), 1))(*[q.q for q in queries])
        self.ctx.check(self.ctx.lib.kg_batch_submit(self.h, qarr, len(queries)))
        self._inflight = (list(queries), qarr, max_groups)

    def poll(self):
        d = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.kg_batch_poll(self.h, ctypes.byref(d)))
        return bool(d.value)

    def collect(self):
        """[(status, codes, values, nulls)] per query, in submission order."""
        queries, _, max_groups = self._inflight
        res, bufs = _result_buffers(queries, max_groups)
        self.ctx.check(self.ctx.lib.kg_batch_collect(self.h, res))
        self._inflight = None
        return _results(queries, res, bufs)

    def ready(self, i):
        """Whether query i of the batch in flight has been published."""
        d = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.kg_batch_query_ready(self.h, i, ctypes.byref(d)))
        return bool(d.value)

    def collect_query(self, i):
        """(status, codes, values, nulls) of query i, once ready(i)."""
        queries, _, max_groups = self._inflight
        res, bufs = _result_buffers([queries[i]], max_groups)
        self.ctx.check(self.ctx.lib.kg_batch_collect_query(self.h, i, res))
        return _results([queries[i]], res, bufs)[0]

    def release(self):
        self.ctx.check(self.ctx.lib.kg_batch_release(self.h))
        self._inflight = None

    def run(self, queries, max_groups=4096):
        self.submit(queries, max_groups)
        return self.collect()

    def free(self):
        if self.h:
            self.ctx.lib.kg_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def serve(ctx, queries, arrivals, mode, slots=3, max_batch=128, arena_bytes=0, max_groups=64, max_batch_rows=0,
          lanes=1, lane_scan_rows=0):
    """Replays `queries` arriving at `arrivals` seconds (ascending). Returns
    (results as Batch.collect, latencies in s, stats dict). Batched mode runs
    what batches do not take on `lanes` plain lanes, and scans of at least
    lane_scan_rows rows arriving with nothing pending on an idle lane."""
    n = len(queries)
    qarr = (abi.KgQuery * max(n, 1))(*[q.q for q in queries])
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    lat = np.zeros(max(n, 1), dtype=np.float64)
    res, bufs = _result_buffers(queries, max_groups)
    opts = abi.KgServeOpts(mode, slots, max_batch, lanes, arena_bytes, max_batch_rows, lane_scan_rows)
    st = abi.KgServeStats()
    ctx.check(ctx.lib.kg_serve(ctx.h, qarr, n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ctypes.byref(opts), res, lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ctypes.byref(st)))
    stats = {k: getattr(st, k) for k, _ in abi.KgServeStats._fields_}
    return _results(queries, res, bufs), lat[:n], stats


# -- ORDER BY and result rows -----------------------------------------------------------
def enum_ranks(strings):
    """Rank of each dictionary string (by code) in the order ENUM values compare."""
    enc = [s.encode() for s in strings]
    off = np.zeros(len(enc) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(b) for b in enc])
    buf = ctypes.create_string_buffer(b"".join(enc) or b"\0")
    out = np.empty(max(len(enc), 1), dtype=np.int32)
    st = abi.load().kg_enum_ranks(buf, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(enc),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if st != abi.KG_OK:
        raise KgError(st, "kg_enum_ranks failed")
    return out[:len(enc)]


def order_by(ctx, keys, ascending, dictionaries=None):
    """Device int32 permutation: input positions in ORDER BY order. keys:
    Column or TextColumn objects of equal length; dictionaries: {key index:
    strings by code} for ENUM keys."""
    n = len(keys[0])
    dictionaries = dictionaries or {}
    dev = f"cuda:{ctx.device}"
    cols = (abi.KgCol * len(keys))()
    texts = (abi.KgTextCol * len(keys))()
    ranks = (ctypes.c_void_p * len(keys))()
    keep = []
    for i, k in enumerate(keys):
        if isinstance(k, TextColumn):
            cols[i] = abi.KgCol(abi.KG_TEXT, 0, 0, 0)
            texts[i] = k.desc()
        else:
            cols[i] = k.desc()
            if k.type == abi.KG_ENUM:
                r = torch.as_tensor(enum_ranks(dictionaries[i])).to(dev)
                keep.append(r)
                ranks[i] = r.data_ptr()
    asc = (ctypes.c_int32 * len(keys))(*[1 if a else 0 for a in ascending])
    perm = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ctx.check(ctx.lib.kg_order_by(ctx.h, n, cols, len(keys), asc, ranks, texts, _ptr(perm)))
    return perm[:n]


def format_rows(columns, order=None):
    """Result rows as the reference prints them. columns: (sql_type, values,
    nulls or None, dictionary) with host numpy values (ENUM: codes and the
    strings by code; TEXT: a list of str)."""
    n = len(columns[0][1]) if columns else 0
    fc = (abi.KgFmtCol * max(len(columns), 1))()
    keep = []
    for i, (typ, vals, nulls, dictionary) in enumerate(columns):
        if typ == abi.KG_TEXT:
            strings = [s.encode() for s in vals]
            off = np.zeros(len(strings) + 1, dtype=np.int64)
            off[1:] = np.cumsum([len(b) for b in strings])
            blob = ctypes.create_string_buffer(b"".join(strings) or b"\0")
            data = np.zeros(1, dtype=np.int8)
            keep += [off, blob, data]
            fc[i] = abi.KgFmtCol(typ, 0, data.ctypes.data, 0, ctypes.addressof(blob), off.ctypes.data)
        else:
            vals = np.ascontiguousarray(vals)
            keep.append(vals)
            d = doff = None
            dn = 0
            if typ == abi.KG_ENUM:
                enc = [s.encode() for s in dictionary]
                doff = np.zeros(len(enc) + 1, dtype=np.int64)
                doff[1:] = np.cumsum([len(b) for b in enc])
                d = ctypes.create_string_buffer(b"".join(enc) or b"\0")
                dn = len(enc)
                keep += [doff, d]
            fc[i] = abi.KgFmtCol(typ, dn, vals.ctypes.data, 0, ctypes.addressof(d) if d is not None else 0,
                                 doff.ctypes.data if doff is not None else 0)
        if nulls is not None:
            nl = np.ascontiguousarray(nulls, dtype=np.int8)
            keep.append(nl)
            fc[i].nulls = nl.ctypes.data
    ordp = 0
    if order is not None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        keep.append(order)
        ordp = order.ctypes.data
    lib = abi.load()
    ln = ctypes.c_int64()
    st = lib.kg_format_rows(fc, len(columns), n, ordp, None, 0, ctypes.byref(ln))
    if st != abi.KG_OK:
        raise KgError(st, "kg_format_rows failed")
    buf = ctypes.create_string_buffer(max(ln.value, 1))
    st = lib.kg_format_rows(fc, len(columns), n, ordp, buf, ln.value, ctypes.byref(ln))
    if st != abi.KG_OK:
        raise KgError(st, "kg_format_rows failed")
    return buf.raw[:ln.value].decode()
