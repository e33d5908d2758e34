"""Build recipe for the kushgpu shared library (gfx950) and the CPU oracle.

    python -m kushdb_amd.build            # libkushgpu.so + oracle + tests/cxx drivers
    python -m kushdb_amd.build --lib-only

The library is compiled in-tree to kushdb_amd/libkushgpu.so with hipcc for
gfx950 only. Packed-FP32 instruction selection is disabled for every HIP
translation unit.
"""
import argparse
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kushdb_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libkushgpu.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

SOURCES = [
    "kg_context.hip",
    "kg_scan.hip",
    "kg_join.hip",
    "kg_index.hip",
    "kg_text.hip",
    "kg_batch.hip",
    "kg_serve.hip",
    "kg_file.hip",
    "kg_order.hip",
    "kg_expr.hip",
    "kg_skinner.hip",
    "kg_group.hip",
    "kg_detsum.hip",
    "kg_chain.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build_lib(verbose=False):
    """Compiles every translation unit (in parallel) and links the library."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *compile_flags, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    cmd = ["make", "-s", "-C", ORACLE_DIR]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_cxx_tests(verbose=False):
    """The C++ drivers of the C ABI under tests/cxx (need the library and the
    oracle built first)."""
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "tests", "cxx")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    build_lib(a.verbose)
    if not a.lib_only:
        build_oracle(a.verbose)
        build_cxx_tests(a.verbose)
    return 0


if __name__ == "__main__":
    sys.exit(main())
