"""Expression trees for the device expression evaluator (kg_expr_eval /
kg_expr_select), built with helpers named after the reference's plan builders
(util/builder.h: ColRef, Literal, Add, Mul, Eq, Lt, And, Or, Not, Case, Float,
Extract, Like, ...). Result types follow the plan's rules
(plan/expression/arithmetic_expression.cc:11-78): comparisons, AND/OR and the
string predicates are BOOLEAN, arithmetic keeps its operand type; And/Or over
a list nest to the left as util/builder.h:194-222 does. Like takes the SQL
pattern ('%', '_'), as the planner's LIKE does before it becomes a regex.
"""
import ctypes

import numpy as np
import torch

from . import abi
from .abi import KgError
from .workloads import julian_day


class Expr:
    def __init__(self, op, typ, kids=(), ilit=0, flit=0.0, null=False, text=None, col=-1, values=None):
        self.op, self.type, self.kids = op, typ, list(kids)
        self.ilit, self.flit, self.null, self.text, self.col, self.values = ilit, flit, null, text, col, values

    def flatten(self):
        """Post-order list of (expr, child indices): children before parents,
        the root last."""
        out = []

        def visit(e):
            idx = [visit(k) for k in e.kids]
            out.append((e, idx))
            return len(out) - 1

        visit(self)
        return out


def ColRef(i, typ):
    """Column i of the column set, of SQL type typ."""
    return Expr(abi.KG_E_COL, typ, col=i)


def Literal(v, typ=None, null=False):
    """A literal: bool (BOOLEAN), float (REAL), str/bytes (TEXT), (y, m, d)
    (DATE), int (INT unless typ says SMALLINT, BIGINT, ENUM (a code), REAL)."""
    if isinstance(v, bool):
        return Expr(abi.KG_E_LIT, abi.KG_BOOLEAN if typ is None else typ, ilit=int(v), null=null)
    if isinstance(v, tuple):
        return Expr(abi.KG_E_LIT, abi.KG_DATE, ilit=julian_day(*v), null=null)
    if isinstance(v, float):
        return Expr(abi.KG_E_LIT, abi.KG_REAL, flit=v, null=null)
    if isinstance(v, (str, bytes)):
        return Expr(abi.KG_E_LIT, abi.KG_TEXT, text=v.encode() if isinstance(v, str) else v, null=null)
    t = abi.KG_INT if typ is None else typ   # KG_SMALLINT is 0
    if t == abi.KG_REAL:
        return Expr(abi.KG_E_LIT, t, flit=float(v), null=null)
    return Expr(abi.KG_E_LIT, t, ilit=int(v), null=null)


def _bin(op, a, b, typ=None):
    return Expr(op, typ if typ is not None else a.type, (a, b))


def Add(a, b): return _bin(abi.KG_E_ADD, a, b)
def Sub(a, b): return _bin(abi.KG_E_SUB, a, b)
def Mul(a, b): return _bin(abi.KG_E_MUL, a, b)
def Div(a, b): return _bin(abi.KG_E_DIV, a, b)
def Eq(a, b): return _bin(abi.KG_E_EQ, a, b, abi.KG_BOOLEAN)
def Neq(a, b): return _bin(abi.KG_E_NEQ, a, b, abi.KG_BOOLEAN)
def Lt(a, b): return _bin(abi.KG_E_LT, a, b, abi.KG_BOOLEAN)
def Leq(a, b): return _bin(abi.KG_E_LEQ, a, b, abi.KG_BOOLEAN)
def Gt(a, b): return _bin(abi.KG_E_GT, a, b, abi.KG_BOOLEAN)
def Geq(a, b): return _bin(abi.KG_E_GEQ, a, b, abi.KG_BOOLEAN)
def StartsWith(a, b): return _bin(abi.KG_E_STARTS_WITH, a, b, abi.KG_BOOLEAN)
def EndsWith(a, b): return _bin(abi.KG_E_ENDS_WITH, a, b, abi.KG_BOOLEAN)
def Contains(a, b): return _bin(abi.KG_E_CONTAINS, a, b, abi.KG_BOOLEAN)


def And(exprs):
    out = _bin(abi.KG_E_AND, exprs[0], exprs[1], abi.KG_BOOLEAN)
    for e in exprs[2:]:
        out = _bin(abi.KG_E_AND, out, e, abi.KG_BOOLEAN)
    return out


def Or(exprs):
    out = _bin(abi.KG_E_OR, exprs[0], exprs[1], abi.KG_BOOLEAN)
    for e in exprs[2:]:
        out = _bin(abi.KG_E_OR, out, e, abi.KG_BOOLEAN)
    return out


def Not(a): return Expr(abi.KG_E_NOT, abi.KG_BOOLEAN, (a,))
def IsNull(a): return Expr(abi.KG_E_IS_NULL, abi.KG_BOOLEAN, (a,))
def Case(cond, then, other): return Expr(abi.KG_E_CASE, then.type, (cond, then, other))
def Float(a): return Expr(abi.KG_E_INT_TO_REAL, abi.KG_REAL, (a,))
def ExtractYear(a): return Expr(abi.KG_E_EXTRACT_YEAR, abi.KG_INT, (a,))


def Like(a, pattern):
    return Expr(abi.KG_E_LIKE, abi.KG_BOOLEAN, (a,), text=pattern.encode() if isinstance(pattern, str) else pattern)


def In(a, codes):
    return Expr(abi.KG_E_IN, abi.KG_BOOLEAN, (a,), values=[int(c) for c in codes])


class Program:
    """ctypes node array of an expression (kept alive with its byte and value
    buffers)."""

    def __init__(self, root):
        flat = root.flatten()
        self.n = len(flat)
        self.nodes = (abi.KgExprNode * self.n)()
        self.keep = []
        for k, (e, idx) in enumerate(flat):
            N = self.nodes[k]
            N.op, N.type = e.op, e.type
            N.a = e.col if e.op == abi.KG_E_COL else (idx[0] if len(idx) > 0 else -1)
            N.b = idx[1] if len(idx) > 1 else -1
            N.c = idx[2] if len(idx) > 2 else -1
            N.flags = abi.KG_E_NULL if e.null else 0
            N.ilit, N.flit = e.ilit, e.flit
            if e.text is not None:
                buf = ctypes.create_string_buffer(e.text or b"\0", max(len(e.text), 1))
                self.keep.append(buf)
                N.str, N.str_len = ctypes.addressof(buf), len(e.text)
            if e.values is not None:
                arr = np.ascontiguousarray(e.values or [0], dtype=np.int32)
                self.keep.append(arr)
                N.values, N.nvalues = arr.ctypes.data, len(e.values)


def _column_set(columns, dictionaries):
    from .runtime import TextColumn
    n = len(columns)
    cols = (abi.KgCol * max(n, 1))()
    texts = (abi.KgTextCol * max(n, 1))()
    dicts = (abi.KgEnumDict * max(n, 1))()
    keep = []
    for i, c in enumerate(columns):
        if isinstance(c, TextColumn):
            cols[i] = abi.KgCol(abi.KG_TEXT, 0, 0, 0)
            texts[i] = c.desc()
        else:
            cols[i] = c.desc()
        d = (dictionaries or {}).get(i)
        if d is not None:
            enc = [s.encode() for s in d]
            off = np.zeros(len(enc) + 1, dtype=np.int64)
            off[1:] = np.cumsum([len(b) for b in enc])
            blob = ctypes.create_string_buffer(b"".join(enc) or b"\0")
            keep += [off, blob]
            dicts[i] = abi.KgEnumDict(ctypes.addressof(blob), off.ctypes.data, len(enc))
    return cols, texts, dicts, keep


_TORCH = {abi.KG_SMALLINT: torch.int16, abi.KG_INT: torch.int32, abi.KG_DATE: torch.int32, abi.KG_ENUM: torch.int32,
          abi.KG_BIGINT: torch.int64, abi.KG_REAL: torch.float64, abi.KG_BOOLEAN: torch.int8}


def expr_eval(ctx, n, columns, expr, rows=None, dictionaries=None):
    """The expression over rows (a device int32 tensor) or 0..n of the column
    set, as a device Column (values and NULL flags)."""
    from .runtime import Column, _ptr
    prog = Program(expr)
    cols, texts, dicts, keep = _column_set(columns, dictionaries)
    m = rows.numel() if rows is not None else n
    dev = f"cuda:{ctx.device}"
    data = torch.empty(max(m, 1), dtype=_TORCH[expr.type], device=dev)
    nulls = torch.empty(max(m, 1), dtype=torch.int8, device=dev)
    out = abi.KgCol(expr.type, 0, data.data_ptr(), nulls.data_ptr())
    ctx.check(ctx.lib.kg_expr_eval(ctx.h, m, cols, texts, dicts, len(columns), prog.nodes, prog.n, _ptr(rows),
                                   ctypes.byref(out)))
    return Column(expr.type, data[:m], nulls[:m])


def expr_select(ctx, n, columns, expr, rows=None, dictionaries=None):
    """Rows (ascending device int32) whose BOOLEAN expression is TRUE and not
    NULL."""
    from .runtime import _ptr
    prog = Program(expr)
    cols, texts, dicts, keep = _column_set(columns, dictionaries)
    m = rows.numel() if rows is not None else n
    out = torch.empty(max(m, 1), dtype=torch.int32, device=f"cuda:{ctx.device}")
    cnt = ctypes.c_int64()
    ctx.check(ctx.lib.kg_expr_select(ctx.h, m, cols, texts, dicts, len(columns), prog.nodes, prog.n, _ptr(rows),
                                     _ptr(out), ctypes.byref(cnt)))
    return out[:cnt.value]


__all__ = ["Expr", "ColRef", "Literal", "Add", "Sub", "Mul", "Div", "Eq", "Neq", "Lt", "Leq", "Gt", "Geq", "And", "Or",
           "Not", "IsNull", "Case", "Float", "ExtractYear", "Like", "In", "StartsWith", "EndsWith", "Contains",
           "Program", "expr_eval", "expr_select", "KgError"]
