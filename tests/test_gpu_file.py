"""On-disk column files (kg_file_*) against the reference's own fixture files:
every typed .kdb / _null.kdb in tests/golden/ref/data loads to the values and
NULL flags numpy reads from the file; the .kdbenum dictionaries give the
strings the fixture tests decode; each .kdbindex equals both the index built
over the loaded column's non-NULL rows and the file's entries as read in
Python; a TEXT column file written in the reference's slot layout
(runtime/column_data.cc:175-229: the fixtures hold no TEXT file) loads back to
its strings; loaded columns run through the operators like any other."""
import os
import struct

import numpy as np
import pytest
import torch

import kdb
from kushdb_amd import abi
from kushdb_amd.runtime import Column, ColumnIndex, Context, text_select

pytestmark = pytest.mark.gpu

TYPES = {"int": abi.KG_INT, "boolean": abi.KG_BOOLEAN, "date": abi.KG_DATE, "real": abi.KG_REAL,
         "smallint": abi.KG_SMALLINT, "bigint": abi.KG_BIGINT, "enum": abi.KG_ENUM}
FILES = {"info_id": "int", "info_cheated": "boolean", "info_date": "date", "info_zscore": "real",
         "info_num1": "smallint", "info_num2": "bigint", "people_id": "int", "n_nationkey": "int",
         "l_discount": "real", "l_extendedprice": "real", "l_quantity": "real", "l_tax": "real",
         "l_shipdate": "date", "l_returnflag_enum": "enum", "l_linestatus_enum": "enum"}
INDEXED = {"info_id": "int", "info_date": "date", "info_num2": "bigint", "info_zscore": "real"}


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.file_release()


def path_of(name, ext=".kdb"):
    return os.path.join(kdb.DATA, name + ext)


@pytest.mark.parametrize("name", sorted(FILES))
def test_typed_column_file(ctx, name):
    typ = FILES[name]
    print("file", name, flush=True)
    col = ctx.file_column(path_of(name), TYPES[typ])
    vals, nulls = kdb.col(name, typ)
    got = col.data.cpu().numpy()
    assert got.dtype == vals.dtype and np.array_equal(got.view(np.uint8), vals.view(np.uint8))
    if nulls is None:
        assert col.nulls is None
    else:
        assert np.array_equal(col.nulls.cpu().numpy(), nulls)
    # cached by path: the same device column again
    again = ctx.file_column(path_of(name), TYPES[typ])
    assert again.data.data_ptr() == col.data.data_ptr()


@pytest.mark.parametrize("name", ["l_returnflag_enum", "l_linestatus_enum"])
def test_enum_dictionary_file(ctx, name):
    assert ctx.file_enum(path_of(name, ".kdbenum")) == kdb.enum_dict(name)


@pytest.mark.parametrize("name", sorted(INDEXED))
def test_index_file(ctx, name):
    typ = INDEXED[name]
    print("index", name, flush=True)
    sql = TYPES[typ]
    col = ctx.file_column(path_of(name), sql)
    idx = ctx.file_index(path_of(name, ".kdbindex"), sql)
    nulls = col.nulls.cpu().numpy() if col.nulls is not None else np.zeros(len(col.data), np.int8)
    live = torch.as_tensor(np.nonzero(nulls == 0)[0].astype(np.int32)).to(col.data.device)
    built = ColumnIndex(ctx, col, live)
    ent = kdb.index_entries(name, typ)
    assert idx.nkeys == built.nkeys == len(ent)
    assert idx.ntids == built.ntids == sum(len(v) for v in ent.values())
    keys = np.array(list(ent.keys()), dtype=col.data.cpu().numpy().dtype)
    kc = Column(sql, torch.as_tensor(keys).to(col.data.device))
    for ix in (idx, built):
        ptrs, cnts = ix.lookup(kc)
        offs = ix.offsets_of(ptrs).cpu().numpy()
        cnts = cnts.cpu().numpy()
        tids = ix.tids().cpu().numpy()
        for k, o, c in zip(keys.tolist(), offs, cnts):
            assert tids[o:o + c].tolist() == ent[k], (name, k)
    idx.free()
    built.free()


def write_text_file(path, strings):
    """runtime/column_data.cc:175-229: int32 cardinality padded to 8 bytes,
    then {int32 length, 4 bytes padding, uint64 offset} per slot, then the
    strings each followed by a NUL byte."""
    head = struct.pack("<i4x", len(strings))
    off = 8 + 16 * len(strings)
    slots, heap = b"", b""
    for s_ in strings:
        b = s_.encode()
        slots += struct.pack("<i4xQ", len(b), off)
        heap += b + b"\0"
        off += len(b) + 1
    with open(path, "wb") as f:
        f.write(head + slots + heap)


def test_text_column_file(ctx, tmp_path):
    strings = ["alpha", "", "beta gamma", "a", "delta" * 20, "aardvark"]
    p = tmp_path / "t_name.kdb"
    write_text_file(p, strings)
    np.array([0, 1, 0, 0, 0, 0], dtype=np.int8).tofile(tmp_path / "t_name_null.kdb")
    t = ctx.file_text(p)
    off = t.offsets.cpu().numpy()
    heap = t.data.cpu().numpy().tobytes()
    assert [heap[off[i]:off[i + 1]].decode() for i in range(len(strings))] == strings
    assert t.nulls.cpu().numpy().tolist() == [0, 1, 0, 0, 0, 0]
    rows = text_select(ctx, t, abi.KG_TEXT_STARTS_WITH, "a")
    assert rows.cpu().numpy().tolist() == [0, 3, 5]
    ctx.file_release(p)


def test_loaded_columns_feed_the_operators(ctx):
    """TPC-H Q6-style selection over loaded lineitem columns equals numpy on
    the file contents."""
    ship = ctx.file_column(path_of("l_shipdate"), abi.KG_DATE)
    disc = ctx.file_column(path_of("l_discount"), abi.KG_REAL)
    s, _ = kdb.col("l_shipdate", "date")
    d, _ = kdb.col("l_discount", "real")
    lo = int(np.quantile(s, 0.25))
    rows = ctx.scan_select([ship, disc], [(0, abi.KG_GEQ, lo), (1, abi.KG_LT, 0.05)])
    want = np.nonzero((s >= lo) & (d < 0.05))[0]
    assert np.array_equal(rows.cpu().numpy(), want)


def test_missing_and_malformed_files(ctx, tmp_path):
    with pytest.raises(RuntimeError):
        ctx.file_column(tmp_path / "absent.kdb", abi.KG_INT)
    bad = tmp_path / "odd.kdb"
    bad.write_bytes(b"\x01\x02\x03")
    with pytest.raises(RuntimeError):
        ctx.file_column(bad, abi.KG_INT)
    with pytest.raises(RuntimeError):
        ctx.file_column(path_of("info_id"), abi.KG_TEXT)
