"""The C++ mesh decoders (include/bvhtracer_mesh.hpp) against the Python
decoders, which the reference's decoder tests and fixtures pin
(tests/test_mesh.py): bit for bit on the reference's test inputs and assets,
and the same errors on malformed input.  Host code only: runs on the CPU."""
import gzip
import io
import json
import os
import subprocess

import numpy as np
import pytest

from bvhtracer_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
EXE = os.path.join(HERE, "cpp", "mesh_dump")


@pytest.fixture(scope="module", autouse=True)
def mesh_dump_built():
    # built by __graft_entry__.build(); host code only, so a checkout that has
    # not run build() yet makes it here (like the oracle library)
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "mesh_dump"], check=True, timeout=300)


def cpp_decode(kind, text, tmp_path):
    path = tmp_path / f"in.{kind}"
    path.write_bytes(text.encode() if isinstance(text, str) else text)
    r = subprocess.run([EXE, kind, str(path)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout


def words(mesh):
    n = mesh.len()
    rows = np.concatenate([mesh.primitives().reshape(n, 9), mesh.tex_coords().reshape(n, 6),
                           mesh.normals().reshape(n, 9)], axis=1).astype(np.float32).view(np.uint32)
    return "".join("".join(f" {w:08x}" for w in row) + "\n" for row in rows)


def golden():
    with open(os.path.join(GOLDEN, "mesh_decoders.json")) as f:
        return json.load(f)


def test_tri_mesh_test_input(tmp_path):
    text = golden()["tri_mesh"]["input"]
    rc, out = cpp_decode("tri", text, tmp_path)
    assert rc == 0 and out == words(M.TriMeshDecoder(io.StringIO(text)).read_mesh())


def test_obj_mesh_test_input(tmp_path):
    text = golden()["obj_mesh"]["input"]
    rc, out = cpp_decode("obj", text, tmp_path)
    assert rc == 0 and out == words(M.ObjMeshDecoder(io.StringIO(text)).read_mesh())


@pytest.mark.parametrize("name", ["cube.obj", "unity.tri.gz", "bigben.tri.gz"])
def test_reference_assets(tmp_path, name):
    raw = open(os.path.join(GOLDEN, name), "rb").read()
    text = gzip.decompress(raw).decode() if name.endswith(".gz") else raw.decode()
    kind = "obj" if ".obj" in name else "tri"
    rc, out = cpp_decode(kind, text, tmp_path)
    assert rc == 0
    assert out == words(M.open_mesh(os.path.join(GOLDEN, name)))


def test_tri_loader_cases(tmp_path):
    for case in golden()["tri_loader"]["cases"]:
        rc, out = cpp_decode("tri", case["input"], tmp_path)
        ref = M.load_tri_text(case["input"])
        assert rc == 0 and out.count("\n") == ref.shape[0], case["name"]
        got = np.array([[int(w, 16) for w in line.split()[:9]] for line in out.splitlines()], np.uint32)
        assert np.array_equal(got.reshape(-1, 9), ref.view(np.uint32).reshape(-1, 9)), case["name"]


@pytest.mark.parametrize("text", ["1 2 3 4 5 6 7 8\n9", "1 2 3 4 5 6 7 8 x", "1 2 3 4 5 6 7 8 0x1p3",
                                  "1 2 3 4 5 6 7 8 1_0", "1 2 3", "1 2 3 4 5 6 7 8 9\r\n"])
def test_tri_loader_errors_match(tmp_path, text):
    rc, out = cpp_decode("tri", text, tmp_path)
    try:
        M.load_tri_text(text)
        py = None
    except M.TriLoaderError as e:
        py = (e.line_number, {"ExpectedFloat": 0, "EndOfFile": 1}[e.kind])
    if py is None:
        assert rc == 0
    else:
        assert rc == 1 and out.startswith(f"TriLoaderError {py[0]} {py[1]} "), out


def test_numbers_round_like_rust(tmp_path):
    """Decimal inputs within 2^-60 of a binary32 midpoint, and overflow."""
    vals = ["0.100000001490116119384765625", "1.00000005960464477539062500000000001",
            "1.000000059604644775390624999999999", "3.4028235677973366e38", "1e39", "-1e-50", "7.006e-46",
            "inf", "-NaN", "+.5e-3"]
    # 3.4028235677973366e38 lies just below FLT_MAX + 2^103, the overflow
    # threshold: it rounds to FLT_MAX, not to infinity
    text = " ".join(vals[:9]) + "\n"
    rc, out = cpp_decode("tri", text, tmp_path)
    ref = M.load_tri_text(text)
    got = np.array([int(w, 16) for w in out.split()[:9]], np.uint32)
    assert rc == 0 and np.array_equal(got, ref.view(np.uint32).reshape(9))
