"""Copies the reference's own end-to-end fixture data (column files and the
expected result tables) used by the parity tests into tests/golden/ref/.

These are data files the reference's tests hold (end_to_end_test/data/*.kdb,
*_null.kdb, *.kdbenum, *.kdbindex and end_to_end_test/<op>/*_expected.tbl);
no reference source is copied. Run in the build container:

    python tests/golden/copy_fixtures.py [/root/reference]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEST = os.path.join(HERE, "ref")

DATA = (
    [f"info_{c}{s}" for c in ("id", "cheated", "date", "zscore", "num1", "num2") for s in (".kdb", "_null.kdb")]
    + ["info_id.kdbindex", "info_num2.kdbindex", "info_date.kdbindex", "info_zscore.kdbindex",
       "people_id.kdb", "people_id_null.kdb", "n_nationkey.kdb",
       "l_shipdate.kdb", "l_quantity.kdb", "l_extendedprice.kdb", "l_discount.kdb", "l_tax.kdb",
       "l_returnflag_enum.kdb", "l_linestatus_enum.kdb",
       "l_returnflag_enum.kdbenum", "l_linestatus_enum.kdbenum",
       "people_name.kdb", "people_name_null.kdb", "people_name_enum.kdb", "people_name_enum_null.kdb",
       "people_name_enum.kdbenum", "c_custkey.kdb", "r_name.kdb", "c_mktsegment.kdb", "o_orderkey.kdb",
       "o_custkey.kdb", "o_orderdate.kdb", "o_shippriority.kdb", "l_orderkey.kdb", "l_shipmode.kdb",
       "l_commitdate.kdb", "l_receiptdate.kdb", "o_orderpriority.kdb", "r_regionkey.kdb", "n_nationkey.kdb",
       "n_regionkey.kdb", "n_name.kdb", "c_nationkey.kdb", "l_suppkey.kdb", "s_suppkey.kdb", "s_nationkey.kdb",
       "l_partkey.kdb", "p_partkey.kdb", "p_type.kdb", "p_size.kdb", "p_brand.kdb", "p_container.kdb",
       "l_shipinstruct.kdb", "c_acctbal.kdb", "c_name.kdb", "c_phone.kdb", "c_address.kdb", "c_comment.kdb",
       "l_returnflag.kdb", "ps_partkey.kdb", "ps_suppkey.kdb", "ps_supplycost.kdb", "ps_availqty.kdb",
       "p_name.kdb", "l_quantity.kdb", "o_totalprice.kdb"]
)

EXPECTED = (
    [f"hash_join/{t}_expected.tbl" for t in ("int", "bigint", "smallint", "date", "real", "boolean", "enum",
                                                "text")]
    + [f"scan_select/{t}_expected.tbl" for t in ("int", "bigint", "date", "real", "smallint", "multiple_eq",
                                                  "boolean_eq", "boolean", "enum", "int_eq", "text")]
    + [f"group_by_aggregate/{t}_expected.tbl" for t in ("int", "bigint", "smallint", "real", "date", "enum", "text")]
    + ["aggregate/int_expected.tbl", "aggregate/empty_expected.tbl", "tpch/q01_expected.tbl"]
    + [f"simd_scan_select/{t}_expected.tbl" for t in ("int", "date_range", "multiple_col", "date", "enum")]
    + [f"order_by/{t}_expected.tbl" for t in ("int", "bigint", "smallint", "date", "real", "enum", "text")]
    + [f"select/{t}_expected.tbl" for t in ("int", "bigint", "smallint", "date", "real", "enum", "text", "boolean")]
    + ["tpch/q06_expected.tbl", "tpch/q03_expected.tbl", "tpch/q12_expected.tbl", "tpch/q05_expected.tbl",
       "tpch/q14_expected.tbl", "tpch/q19_expected.tbl", "tpch/q18_expected.tbl",
       "tpch/q10_expected.tbl", "tpch/q11_expected.tbl", "tpch/q07_expected.tbl", "tpch/q09_expected.tbl",
       "cross_product/cross_product_expected.tbl"]
    + [f"skinner_join/{t}_expected.tbl" for t in ("int", "smallint", "date", "bigint", "real", "boolean",
                                                   "enum", "text", "multiple_tables_join")]
)


def main(ref="/root/reference"):
    e2e = os.path.join(ref, "end_to_end_test")
    os.makedirs(os.path.join(DEST, "data"), exist_ok=True)
    for f in DATA:
        shutil.copyfile(os.path.join(e2e, "data", f), os.path.join(DEST, "data", f))
    for f in EXPECTED:
        d = os.path.join(DEST, os.path.dirname(f))
        os.makedirs(d, exist_ok=True)
        shutil.copyfile(os.path.join(e2e, f), os.path.join(DEST, f))


if __name__ == "__main__":
    main(*sys.argv[1:])
