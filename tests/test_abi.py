"""The C ABI library loads and exports every entry point include/bvht.h
declares (no compute calls: these run without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bvht.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bvht_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bvht_trace_blas" in syms and "bvht_blas_build" in syms and len(syms) >= 30


def test_library_exports_every_declared_symbol():
    import bvhtracer_amd._lib as L

    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libbvht_hip.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (bvht_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = L.lib()  # ctypes binding of every signature resolves
    assert set(L.SIGNATURES) == set(declared_symbols())
    assert lib.bvht_last_error() is not None


def header_arities():
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "bvht.h")).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bvht_\w+)\s*\(([^;]*?)\)\s*;", text, re.M | re.S):
        args = " ".join(args.split())
        out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_rust_sys_crate_declares_every_entry_point():
    """No Rust toolchain exists here, so the -sys crate is checked against
    the header as text: the same entry points with the same number of
    arguments, and the record sizes it asserts."""
    src = open(os.path.join(ROOT, "rust", "bvhtracer-hip-sys", "src", "lib.rs")).read()
    rust = {}
    for name, args in re.findall(r"pub fn (bvht_\w+)\((.*?)\)", src, re.S):
        args = " ".join(args.split())
        rust[name] = 0 if not args else args.count(",") + 1
    assert rust == header_arities()
    for rec, size in (("bvht_node", 32), ("bvht_tnode", 32), ("bvht_ray", 32), ("bvht_hit", 24),
                      ("bvht_instance", 144), ("bvht_camera", 100), ("bvht_body", 272), ("bvht_force", 48)):
        assert f"size_of::<{rec}>() == {size})" in src, rec
