"""The oracle's TEXT restatement against Python's own string operations and
re.fullmatch of the planner's escaped LIKE regex (plan/planner.cc:284-323)."""
import ctypes
import re

import numpy as np
import pytest

from oracle import pyoracle as po

EQ, NEQ, LT, LEQ, GT, GEQ, STARTS, ENDS, CONTAINS, LIKE = range(10)


def pred(op, s, p):
    sb, pb = s.encode(), p.encode()
    return bool(po.lib().kor_text_pred(op, sb, len(sb), pb, len(pb)))


def like_regex(pattern):
    """The planner's conversion: escape regex metacharacters, '_' -> '.', '%' -> '.*', anchored."""
    out = "^"
    for c in pattern:
        if c in ".+*?^$()[]{}|\\":
            out += "\\" + c
        elif c == "_":
            out += "."
        elif c == "%":
            out += ".*"
        else:
            out += c
    return out + "$"


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
    rng = np.random.default_rng(2)
    alpha = list("abc.*+?()[]")
    palpha = alpha + ["%", "_", "%"]
    for _ in range(5000):
        s = "".join(rng.choice(alpha, rng.integers(0, 9)))
        p = "".join(rng.choice(palpha, rng.integers(0, 6)))
        want = re.fullmatch(like_regex(p), s, flags=re.DOTALL) is not None
        assert pred(LIKE, s, p) == want, (s, p)


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
