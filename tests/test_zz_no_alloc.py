"""kg_scan_select_noalloc (the scan-select entry point that allocates nothing,
its tile counts and offsets in the library's static per-GPU scratch array)
against kg_scan_select through the C ABI: the same selected rows, in order,
on two calls in a row (the second over fewer rows, so the scratch keeps the
first call's longer contents), on an empty input, and KG_ERR_CAPACITY beyond
its 536,870,912-row limit."""
import ctypes

import numpy as np
import pytest
import torch

from kushdb_amd import abi
from kushdb_amd.runtime import Column, Context, _pred

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _select(ctx, fn, cols, preds, conj=1, n=None):
    n = len(cols[0]) if n is None else n
    carr = (abi.KgCol * len(cols))(*[c.desc() for c in cols])
    parr = (abi.KgPred * max(len(preds), 1))(*[_pred(cols, p) for p in preds])
    out = torch.full((max(n, 1),), -7, dtype=torch.int32, device=DEV)
    cnt = ctypes.c_int64(-1)
    st = getattr(ctx.lib, fn)(ctx.h, n, carr, len(cols), parr, len(preds), conj, ctypes.c_void_p(out.data_ptr()),
                              ctypes.byref(cnt))
    return st, out[: max(cnt.value, 0)].cpu().numpy(), cnt.value


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_noalloc_equals_scan_select_twice(ctx):
    rng = np.random.default_rng(5)
    for n, conj in ((3_000_000, 1), (700_001, 0)):
        print("noalloc select", n, conj, flush=True)
        day = Column(abi.KG_DATE, torch.as_tensor(rng.integers(0, 2557, n).astype(np.int32)).to(DEV))
        v = Column(abi.KG_REAL, torch.as_tensor(rng.uniform(0, 1, n)).to(DEV),
                   torch.as_tensor((rng.random(n) < 0.1).astype(np.int8)).to(DEV))
        preds = [(0, abi.KG_LEQ, 1800), (1, abi.KG_GT, 0.25)]
        st0, want, c0 = _select(ctx, "kg_scan_select", [day, v], preds, conj)
        st1, got, c1 = _select(ctx, "kg_scan_select_noalloc", [day, v], preds, conj)
        assert st0 == abi.KG_OK and st1 == abi.KG_OK
        assert c1 == c0 and np.array_equal(got, want)


def test_noalloc_empty_and_limit(ctx):
    x = Column(abi.KG_INT, torch.arange(10, dtype=torch.int32, device=DEV))
    st, got, c = _select(ctx, "kg_scan_select_noalloc", [x], [(0, abi.KG_LT, 5)], n=0)
    assert st == abi.KG_OK and c == 0
    limit = 131072 * 4096
    big = Column(abi.KG_BOOLEAN, torch.zeros(limit + 1, dtype=torch.int8, device=DEV))
    st, _, _ = _select(ctx, "kg_scan_select_noalloc", [big], [(0, abi.KG_EQ, 1)])
    assert st == abi.KG_ERR_CAPACITY
    assert b"scratch limit" in ctx.lib.kg_last_error(ctx.h)
