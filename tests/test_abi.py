"""The C-ABI library loads and exports every function include/kushgpu.h
declares (no compute calls: runs without a GPU)."""
import os
import re

from kushdb_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(kg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "kg_ctx_create" in names and "kg_ht_probe" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert declared_functions() <= set(abi.SIGNATURES), sorted(declared_functions() - set(abi.SIGNATURES))


def test_errors_without_context():
    lib = abi.load()
    assert lib.kg_ctx_destroy(None) == abi.KG_OK
    assert lib.kg_last_error(None) == b"null context"
    assert lib.kg_scan_select(None, 0, None, 0, None, 0, 1, None, None) == abi.KG_ERR_INVALID
    assert lib.kg_ht_build(None, None, None, None, 0) == abi.KG_ERR_INVALID
