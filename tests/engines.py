"""Engines the fixture queries (queries.py) run on.

OracleEngine: the CPU restatement (oracle/liboracle.so) — the checker.
GpuEngine:    the product path, libkushgpu.so through its C ABI.
"""
import numpy as np

from oracle import pyoracle as po

_PO_TYPE = {"smallint": po.SMALLINT, "int": po.INT, "bigint": po.BIGINT, "real": po.REAL, "date": po.DATE,
            "boolean": po.BOOLEAN, "enum": po.ENUM}


class OracleEngine:
    name = "oracle"

    def _col(self, c):
        typ, v, n = c
        return po.Col(_PO_TYPE[typ], v, n)

    def select(self, cols, preds, conj):
        return po.scan_select([self._col(c) for c in cols], preds, conj)

    def join(self, b, p):
        return po.hash_join([self._col(b)], [self._col(p)])

    def join_text(self, b, p):
        return po.text_hash_join(b[0], b[1], p[0], p[1])

    def group_agg(self, codes, card, aggs, rows, keys_out=False):
        keys = [po.Col(po.INT, codes)]
        pa = [(k, self._col(c) if c is not None else None) for k, c in aggs]
        kv, _, vals, nulls = po.group_by(keys, pa, rows=rows, max_groups=4096)
        if keys_out:
            return [(int(kv[g, 0]), vals[g], nulls[g]) for g in range(len(vals))]
        return [(vals[g], nulls[g]) for g in range(len(vals))]

    def aggregate(self, aggs, n, rows):
        pa = [(k, self._col(c) if c is not None else None) for k, c in aggs]
        got, v, nl = po.aggregate(pa, n, rows)
        return (v, nl) if got else None

    def skinner(self, cards, conds, **kw):
        """conds: [(lt, (typ, values, nulls), rt, (typ, values, nulls))]."""
        return po.skinner_join(cards, [(lt, lc[1], lc[2], rt, rc[1], rc[2]) for lt, lc, rt, rc in conds])


class GpuEngine:
    name = "gpu"

    def __init__(self):
        import torch
        from kushdb_amd import abi
        from kushdb_amd.runtime import Column, Context, HashTable
        self.torch = torch
        self.abi = abi
        self.Column = Column
        self.HashTable = HashTable
        self.ctx = Context(0)
        self.dev = "cuda:0"
        self._t = {"smallint": abi.KG_SMALLINT, "int": abi.KG_INT, "bigint": abi.KG_BIGINT, "real": abi.KG_REAL,
                   "date": abi.KG_DATE, "boolean": abi.KG_BOOLEAN, "enum": abi.KG_ENUM}

    def _col(self, c):
        typ, v, n = c
        t = self.torch
        data = t.from_numpy(np.ascontiguousarray(v)).to(self.dev)
        nulls = t.from_numpy(np.ascontiguousarray(n, dtype=np.int8)).to(self.dev) if n is not None else None
        return self.Column(self._t[typ], data, nulls)

    def select(self, cols, preds, conj):
        rows = self.ctx.scan_select([self._col(c) for c in cols], preds, conj)
        return rows.cpu().numpy()

    def _text(self, strings, nulls):
        from kushdb_amd.runtime import TextColumn
        return TextColumn.from_strings(strings, device=self.dev,
                                       nulls=None if nulls is None else np.ascontiguousarray(nulls, dtype=np.int8))

    def skinner(self, cards, conds, budget=0, seed=1):
        """TEXT conditions join on device string hashes (kg_text_hash) and the
        result tuples are then checked string by string (kg_text_pairs_equal)."""
        from kushdb_amd.runtime import skinner_join, text_hash, text_pairs_equal
        cache = {}

        def dev(c):
            if id(c[1]) not in cache:
                if c[0] == "text":
                    tc = self._text(c[1], c[2])
                    h, nl = text_hash(self.ctx, tc)
                    cache[id(c[1])] = (self.Column(self.abi.KG_BIGINT, h, nl), tc)
                else:
                    cache[id(c[1])] = (self._col(c), None)
            return cache[id(c[1])]
        t, self.last_stats = skinner_join(self.ctx, [(c, None) for c in cards],
                                          [(lt, dev(lc)[0], rt, dev(rc)[0]) for lt, lc, rt, rc in conds],
                                          budget=budget, seed=seed)
        for lt, lc, rt, rc in conds:
            if lc[0] == "text" and t.shape[0]:
                keep = text_pairs_equal(self.ctx, dev(lc)[1], t[:, lt].contiguous(), dev(rc)[1], t[:, rt].contiguous())
                t = t[keep.long()]
        return sorted(map(tuple, t.cpu().numpy().tolist()))

    def join_text(self, b, p):
        """TEXT equi-join: BIGINT join on kg_text_hash keys, then the exact
        pair check (kg_text_pairs_equal)."""
        from kushdb_amd.runtime import text_hash, text_pairs_equal
        bt, pt = self._text(*b), self._text(*p)
        bh, bn = text_hash(self.ctx, bt)
        ph, pn = text_hash(self.ctx, pt)
        ht = self.HashTable(self.ctx, max(len(bt), 1))
        ht.insert(self.Column(self.abi.KG_BIGINT, bh, bn))
        pr, br = ht.probe(self.Column(self.abi.KG_BIGINT, ph, pn))
        ht.free()
        keep = text_pairs_equal(self.ctx, bt, br, pt, pr).long()
        return pr[keep].cpu().numpy(), br[keep].cpu().numpy()

    def join(self, b, p):
        bc, pc = self._col(b), self._col(p)
        ht = self.HashTable(self.ctx, max(len(bc), 1))
        ht.insert(bc)
        pr, br = ht.probe(pc)
        out = (pr.cpu().numpy(), br.cpu().numpy())
        ht.free()
        return out

    def group_agg(self, codes, card, aggs, rows, keys_out=False):
        return self._agg(codes, card, aggs, rows, keys_out)

    def _agg(self, codes, card, aggs, rows, keys_out):
        abi = self.abi
        n = len(codes) if codes is not None else len(next(c for _, c in aggs if c is not None)[1])
        cols = []
        if codes is not None:
            cols.append(self._col(("int", codes, None)))
        # Row selection as a predicate on a row-id mask column.
        preds = []
        if rows is not None:
            mask = np.zeros(n, dtype=np.int32)
            mask[rows] = 1
            cols.append(self._col(("int", mask, None)))
            preds.append((len(cols) - 1, abi.KG_EQ, 1))
        al = []
        for k, c in aggs:
            if c is None:
                al.append((k, -1))
            else:
                cols.append(self._col(c))
                al.append((k, len(cols) - 1))
        keys = [(0, card)] if codes is not None else []
        kc, vals, nulls = self.ctx.scan_group_agg(cols, preds, keys, al)
        if codes is None:
            return (vals[0], nulls[0]) if len(vals) else None
        if keys_out:
            return [(int(kc[g, 0]), vals[g], nulls[g]) for g in range(len(vals))]
        return [(vals[g], nulls[g]) for g in range(len(vals))]

    def aggregate(self, aggs, n, rows):
        return self._agg(None, 1, aggs, rows, False)
