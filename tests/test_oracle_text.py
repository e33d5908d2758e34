"""The oracle's TEXT restatement against Python's own string operations and
re.fullmatch of the planner's escaped LIKE regex (plan/planner.cc:284-323).

The reference evaluates that regex with RE2::FullMatch and default options
(runtime/string.cc:79-82): the text is UTF-8 and "." is one character other
than '\n'. Python's re on str has the same "." without DOTALL (one code point,
never '\n'), so valid UTF-8 strings with newlines and 2-4-byte characters are
checked against it. For byte strings that are not valid UTF-8, the check is a
bytes regex whose "." is re2's compiled byte ranges for [^\n]."""
import ctypes
import re

import numpy as np
import pytest

from oracle import pyoracle as po

EQ, NEQ, LT, LEQ, GT, GEQ, STARTS, ENDS, CONTAINS, LIKE = range(10)


def pred(op, s, p):
    sb, pb = s.encode(), p.encode()
    return bool(po.lib().kor_text_pred(op, sb, len(sb), pb, len(pb)))


def like_regex(pattern, dot="."):
    """The planner's conversion: escape regex metacharacters, '_' -> '.', '%' -> '.*',
    anchored (\\A / \\Z: Python's '$' would also match before a final newline)."""
    out = r"\A"
    for c in pattern:
        if c in ".+*?^$()[]{}|\\":
            out += "\\" + c
        elif c == "_":
            out += dot
        elif c == "%":
            out += dot + "*"
        else:
            out += c
    return out + r"\Z"


# re2's "." over UTF-8 bytes (code points 00-09, 0B-10FFFF; the byte ranges it
# compiles for 80-10FFFF accept any continuation bytes after C2-DF, E0-EF, F0-F4).
RE2_DOT = rb"(?:[\x00-\x09\x0b-\x7f]|[\xc2-\xdf][\x80-\xbf]|[\xe0-\xef][\x80-\xbf]{2}|[\xf0-\xf4][\x80-\xbf]{3})"


def like_regex_bytes(pattern):
    out = rb"\A"
    for c in pattern:
        b = bytes([c])
        if b in b".+*?^$()[]{}|\\":
            out += b"\\" + b
        elif b == b"_":
            out += RE2_DOT
        elif b == b"%":
            out += RE2_DOT + b"*"
        else:
            out += re.escape(b)
    return out + rb"\Z"


def pred_bytes(op, sb, pb):
    return bool(po.lib().kor_text_pred(op, sb, len(sb), pb, len(pb)))


def test_simple_predicates_vs_python():
    rng = np.random.default_rng(1)
    alpha = "ab.%_"
    for _ in range(3000):
        s = "".join(rng.choice(list(alpha), rng.integers(0, 7)))
        p = "".join(rng.choice(list(alpha), rng.integers(0, 4)))
        assert pred(STARTS, s, p) == s.startswith(p)
        assert pred(ENDS, s, p) == s.endswith(p)
        assert pred(CONTAINS, s, p) == (p in s)
        assert pred(EQ, s, p) == (s == p)
        assert pred(LT, s, p) == (s.encode() < p.encode())
        assert pred(GEQ, s, p) == (s.encode() >= p.encode())


def test_like_vs_regex():
    """ASCII metacharacters, newlines and multi-byte characters in text and
    pattern, against Python's re on str without DOTALL."""
    rng = np.random.default_rng(2)
    alpha = list("abc.*+?()[]") + ["\n", "\u00e9", "\u20ac", "\U0001f600"]
    palpha = alpha + ["%", "_", "%", "_"]
    for _ in range(8000):
        s = "".join(rng.choice(alpha, rng.integers(0, 9)))
        p = "".join(rng.choice(palpha, rng.integers(0, 6)))
        want = re.fullmatch(like_regex(p), s) is not None
        assert pred(LIKE, s, p) == want, (s, p)


def test_like_known_answers():
    """Cases that separate re2's "." from a byte wildcard."""
    assert pred(LIKE, "\u00e9", "_")             # one 2-byte character
    assert not pred(LIKE, "\u00e9", "__")        # not two bytes
    assert pred(LIKE, "a\U0001f600b", "a_b")     # one 4-byte character
    assert not pred(LIKE, "a\nb", "a_b")         # '.' never matches '\n'
    assert not pred(LIKE, "a\nb", "a%b")         # nor does '.*'
    assert pred(LIKE, "a\nb", "a\nb")           # a literal newline does
    assert pred(LIKE, "x\ny", "x\n%")
    assert not pred(LIKE, "ab\n", "a%")          # no '$' before a final newline
    assert pred(LIKE, "", "%")
    assert not pred(LIKE, "", "_")


def test_like_invalid_utf8():
    """Text bytes that are not UTF-8 are matched only by literals (re2's "."
    does not take them); a pattern that is not UTF-8 is rejected by re2 and
    matches nothing."""
    rng = np.random.default_rng(4)
    atoms = [b"a", b"\n", b"\xc3\xa9", b"\x80", b"\xc3", b"\xff", b"\xe2\x82\xac", b"\xed\xa0\x80", b"\xe0\x80\x80"]
    patoms = [b"a", b"\n", b"\xc3\xa9", b"%", b"_", b"%", b"_", b"\xe2\x82\xac"]
    for _ in range(8000):
        s = b"".join(atoms[i] for i in rng.integers(0, len(atoms), rng.integers(0, 7)))
        p = b"".join(patoms[i] for i in rng.integers(0, len(patoms), rng.integers(0, 5)))
        want = re.fullmatch(like_regex_bytes(p), s) is not None
        assert pred_bytes(LIKE, s, p) == want, (s, p)
    for bad in (b"\x80", b"a\xc3", b"\xc0\xaf", b"\xe0\x80\x80", b"\xf4\x90\x80\x80", b"\xf8"):
        for s in (b"", b"a", bad, b"a" + bad):
            assert not pred_bytes(LIKE, s, bad)
            assert not pred_bytes(LIKE, s, b"%" + bad + b"%")


def test_minmax_vs_python():
    rng = np.random.default_rng(3)
    words = ["".join(rng.choice(list("abcdz"), rng.integers(0, 6))) for _ in range(500)]
    enc = [w.encode() for w in words]
    off = np.zeros(len(enc) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(b) for b in enc])
    heap = np.frombuffer(b"".join(enc) + b"\0", dtype=np.uint8)
    nulls = (rng.random(len(enc)) < 0.2).astype(np.int8)
    for is_max in (0, 1):
        pos = po.lib().kor_text_minmax(len(enc), off.ctypes.data, heap.ctypes.data, nulls.ctypes.data, None, is_max)
        live = [b for b, nl in zip(enc, nulls) if not nl]
        assert enc[pos] == (max(live) if is_max else min(live))
        first = next(i for i, (b, nl) in enumerate(zip(enc, nulls)) if not nl and b == enc[pos])
        assert pos == first
