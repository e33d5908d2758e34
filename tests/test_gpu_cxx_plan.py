"""The small star query driven from C++ through the C ABI with the reference's
pipeline lifetimes (init / body / reset per pipeline, a predecessor reset after
its last successor; execution/executable_query.cc:70-95): tests/cxx/c0_plan.cc,
three runs on one context, each checked against the oracle's kor_q_star
(categories and counts exact, sums within 1e-12 relative)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cxx", "c0_plan")


def test_c0_plan_through_the_abi():
    assert os.path.exists(EXE), "tests/cxx/c0_plan is not built (python -m kushdb_amd.build)"
    print("running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "c0 plan ok"
    assert r.stdout.count("groups equal to the oracle") == 3
