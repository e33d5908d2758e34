"""LIKE on the device with the reference's re2 semantics, against the oracle.

The reference evaluates LIKE as RE2::FullMatch of the planner's anchored regex
(plan/planner.cc:284-323, runtime/string.cc:79-82) with default options: '_'
is one UTF-8 character and never '\\n', '%' any run of such characters, and a
pattern that is not UTF-8 matches nothing. Both device entry points that take
LIKE — kg_text_select and kg_expr_select (TEXT and ENUM operands) — are
compared row for row with the oracle on strings with newlines, 2-4-byte
characters and bytes that are not UTF-8."""
import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.expr import ColRef, Like, expr_select
from kushdb_amd.runtime import Column, Context, TextColumn, text_select
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ATOMS = [b"a", b"b", b"e", b"\n", b"\xc3\xa9", b"\xe2\x82\xac", b"\xf0\x9f\x98\x80", b"\x80", b"\xc3", b"\xff"]
PATTERNS = [b"%", b"_", b"a_b", b"a%b", b"%e_a%", b"%\xc3\xa9%", b"_\xe2\x82\xac%", b"%\n%", b"a\nb", b"__",
            b"%_", b"a%", b"%a", b"", b"\x80", b"%\xc3%", b"e%e", b"%b_"]


def _strings(n, seed):
    rng = np.random.default_rng(seed)
    return [b"".join(ATOMS[i] for i in rng.integers(0, len(ATOMS), rng.integers(0, 9))) for _ in range(n)]


def _oracle_col(strs, nulls=None):
    off = np.zeros(len(strs) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(s) for s in strs])
    return po.TextCol(off, np.frombuffer(b"".join(strs) or b"\0", dtype=np.uint8), nulls)


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


@pytest.mark.parametrize("pattern", PATTERNS)
def test_text_select_like(ctx, pattern):
    print("LIKE", pattern, flush=True)
    strs = _strings(20000, 11)
    nulls = (np.random.default_rng(12).random(len(strs)) < 0.1).astype(np.int8)
    col = TextColumn.from_strings(strs, nulls=nulls)
    got = text_select(ctx, col, abi.KG_TEXT_LIKE, pattern).cpu().numpy()
    want = po.text_select(_oracle_col(strs, nulls), abi.KG_TEXT_LIKE, pattern)
    assert np.array_equal(got, want)
    rows = torch.arange(0, len(strs), 3, dtype=torch.int32, device="cuda:0")
    got = text_select(ctx, col, abi.KG_TEXT_LIKE, pattern, rows).cpu().numpy()
    want = po.text_select(_oracle_col(strs, nulls), abi.KG_TEXT_LIKE, pattern, rows.cpu().numpy())
    assert np.array_equal(got, want)


def test_expr_like_text_and_enum(ctx):
    strs = _strings(5000, 13)
    words = ["aéb", "a\nb", "ab", "€", "e\U0001f600a", "a", "", "eea"]
    codes = np.random.default_rng(14).integers(0, len(words), len(strs)).astype(np.int32)
    tcol = TextColumn.from_strings(strs)
    ecol = Column(abi.KG_ENUM, torch.as_tensor(codes).to("cuda:0"))
    oc = _oracle_col(strs)
    for pat in PATTERNS:
        print("expr LIKE", pat, flush=True)
        got = expr_select(ctx, len(strs), [tcol], Like(ColRef(0, abi.KG_TEXT), pat)).cpu().numpy()
        want = po.text_select(oc, abi.KG_TEXT_LIKE, pat)
        assert np.array_equal(got, want), pat
        got = expr_select(ctx, len(strs), [ecol], Like(ColRef(0, abi.KG_ENUM), pat), dictionaries={0: words})
        wenc = [w.encode() for w in words]
        hit = np.array([po.text_select(_oracle_col([w]), abi.KG_TEXT_LIKE, pat).size == 1 for w in wenc])
        assert np.array_equal(got.cpu().numpy(), np.nonzero(hit[codes])[0]), pat


def test_text_hash_verify(ctx):
    """kg_text_hash_verify counts the values whose string differs from the
    string of the smallest row with the same code: 0 for kg_text_hash's own
    codes (duplicates, NULLs, empty strings, a row subset), and for codes that
    collide on purpose (the string's length) exactly the count restated from
    that definition in Python, on all rows and on a subset."""
    from kushdb_amd.runtime import text_hash, text_hash_verify
    rng = np.random.default_rng(31)
    base = _strings(300, 5) + [b"", b"x", "é".encode(), b"a\nb"]
    strs = [base[int(i)] for i in rng.integers(0, len(base), 20000)]
    nulls = (rng.random(len(strs)) < 0.05).astype(np.int8)
    col = TextColumn.from_strings(strs, nulls=nulls)
    dev = "cuda:0"
    sub = torch.as_tensor(np.sort(rng.choice(len(strs), 7000, replace=False)).astype(np.int32)).to(dev)
    for rows in (None, sub):
        codes, _ = text_hash(ctx, col, rows)
        assert text_hash_verify(ctx, col, codes, rows) == 0
        idx = np.arange(len(strs)) if rows is None else sub.cpu().numpy()
        lens = np.array([len(strs[i]) for i in idx], dtype=np.int64)
        fake = torch.as_tensor(lens).to(dev)
        rep = {}
        for i, ln in zip(idx, lens):
            if not nulls[i]:
                rep[ln] = min(rep.get(ln, i), i)
        want = sum(1 for i, ln in zip(idx, lens) if not nulls[i] and strs[i] != strs[rep[ln]])
        assert want > 0
        assert text_hash_verify(ctx, col, fake, rows) == want
    assert text_hash_verify(ctx, col, torch.empty(0, dtype=torch.int64, device=dev), None) == 0
