"""The Python reader of the reference's .kdbindex files (tests/kdb.py), which
the GPU loader tests use as their second check, pinned on the CPU against the
fixture columns themselves: every non-NULL row is listed exactly once, under
its own value, and tuple ids ascend within a key (rows are appended in load
order, util/load.h:40-68)."""
import numpy as np
import pytest

import kdb

INDEXED = {"info_id": "int", "info_date": "date", "info_num2": "bigint", "info_zscore": "real"}


@pytest.mark.parametrize("name", sorted(INDEXED))
def test_index_file_matches_column(name):
    typ = INDEXED[name]
    vals, nulls = kdb.col(name, typ)
    ent = kdb.index_entries(name, typ)
    listed = sorted(t for tids in ent.values() for t in tids)
    live = [i for i in range(len(vals)) if nulls is None or not nulls[i]]
    assert listed == live
    for key, tids in ent.items():
        assert tids == sorted(tids)
        for t in tids:
            assert vals[t] == key
