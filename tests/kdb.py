"""Readers for the reference's column-file fixtures and the result-table text
format, used by the parity tests.

Formats (reference):
  * .kdb typed columns: a plain array of the column's C type
    (runtime/column_data.cc:24-116); *_null.kdb: int8, 1 = NULL.
  * .kdbenum dictionary: {u64 entry_offset; u64 mask; u64 array[]} followed
    by EnumEntry {u64 str_offset; u32 str_len; u32 value} indexed by code
    (runtime/enum.cc:26-36, 38-131).
  * result rows: every value followed by '|' (runtime/printer.cc:13-35); REAL
    with std::fixed (6 decimals), DATE as YYYY-MM-DD, BOOLEAN as t/f.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "ref")
DATA = os.path.join(REF, "data")

_DT = {"smallint": np.int16, "int": np.int32, "bigint": np.int64, "real": np.float64, "date": np.int32,
       "boolean": np.int8, "enum": np.int32}


def col(name, typ):
    """Returns (values, nulls or None) for end_to_end_test/data/<name>.kdb."""
    v = np.fromfile(os.path.join(DATA, name + ".kdb"), dtype=_DT[typ])
    npath = os.path.join(DATA, name + "_null.kdb")
    nulls = np.fromfile(npath, dtype=np.int8) if os.path.exists(npath) else None
    return v, nulls


def text(name):
    """Returns (strings, nulls or None) of a TEXT column file: {int32 count;
    4 bytes padding} then {int32 length; 4 bytes padding; uint64 offset} per
    row, offsets from the file start (runtime/column_data.cc:175-229)."""
    raw = open(os.path.join(DATA, name + ".kdb"), "rb").read()
    n = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=0)[0])
    slots = np.frombuffer(raw, dtype=np.dtype([("len", "<i4"), ("pad", "<i4"), ("off", "<u8")]), count=n, offset=8)
    out = [raw[int(o):int(o) + int(ln)].decode() for ln, o in zip(slots["len"], slots["off"])]
    npath = os.path.join(DATA, name + "_null.kdb")
    nulls = np.fromfile(npath, dtype=np.int8) if os.path.exists(npath) else None
    return out, nulls


def enum_dict(name):
    """Code -> string of a .kdbenum file (reads the entry array by code)."""
    raw = open(os.path.join(DATA, name + ".kdbenum"), "rb").read()
    entry_offset = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=0)[0])
    # Entries are written by code, followed by the string heap; the first
    # entry's string offset marks the end of the entry array.
    first_str = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=entry_offset)[0])
    out = []
    for i in range((first_str - entry_offset) // 16):
        off = entry_offset + 16 * i
        str_offset = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=off)[0])
        str_len = int(np.frombuffer(raw, dtype=np.uint32, count=1, offset=off + 8)[0])
        value = int(np.frombuffer(raw, dtype=np.uint32, count=1, offset=off + 12)[0])
        assert value == i
        out.append(raw[str_offset:str_offset + str_len].decode())
    return out


def index_entries(name, typ):
    """key -> tuple ids of end_to_end_test/data/<name>.kdbindex: {u64 mask;
    u64 array[mask + 1]} then entries {key padded to 8 B; u64 size; u64 next;
    int32 values[size]} back to back (runtime/column_index.cc:150-240)."""
    raw = open(os.path.join(DATA, name + ".kdbindex"), "rb").read()
    mask = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=0)[0])
    off = 8 + 8 * (mask + 1)
    dt = np.dtype(_DT[typ])
    out = {}
    while off < len(raw):
        key = np.frombuffer(raw, dtype=dt, count=1, offset=off)[0].item()
        size = int(np.frombuffer(raw, dtype=np.uint64, count=1, offset=off + 8)[0])
        out[key] = np.frombuffer(raw, dtype=np.int32, count=size, offset=off + 24).tolist()
        off += 24 + 4 * size
    assert off == len(raw)
    return out


def expected(rel):
    with open(os.path.join(REF, rel)) as f:
        return [ln for ln in f.read().split("\n") if ln != ""]


def fmt_date(jd):
    from oracle import pyoracle as po
    y, m, d = po.split_date(int(jd))
    return f"{y}-{m:02d}-{d:02d}"


def fmt(typ, v, null=False):
    """One printed value (without the trailing '|')."""
    if null:
        return ""
    if typ == "real":
        return f"{float(v):.6f}"
    if typ == "date":
        return fmt_date(v)
    if typ == "boolean":
        return "t" if v else "f"
    return str(int(v))


def row(vals):
    return "".join(f"{v}|" for v in vals)


def bits_f64(b):
    return float(np.array([b], dtype=np.int64).view(np.float64)[0])
