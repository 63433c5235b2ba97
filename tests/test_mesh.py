"""Mesh ingestion against the reference's decoder tests and assets
(tests/golden/mesh_decoders.json, unity.tri.gz, bigben.tri.gz, cube.obj):
bit-exact triangles, texture coordinates and face normals."""
import io
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from bvhtracer_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = json.load(open(os.path.join(GOLDEN, "mesh_decoders.json")))


def f32(nested):
    flat = [x for tri in nested for v in tri for x in v]
    return M.parse_f32(flat) if flat else np.zeros(0, np.float32)


def test_tri_mesh_decoder():
    """test_tri_mesh.rs: triangles, zero texture coordinates, and face
    normals bit for bit (the second triangle's normal
    (-0.86602545, 0.49999988, -1.7462564e-7) pins normalize and cross)."""
    fx = FIX["tri_mesh"]
    m = M.TriMeshDecoder(io.StringIO(fx["input"])).read_mesh()
    assert m.len() == 2
    assert m.primitives().tobytes() == f32(fx["primitives"]).tobytes()
    assert m.tex_coords().tobytes() == f32(fx["tex_coords"]).tobytes()
    # compared as floats, as the reference's assert_eq! does: the first
    # normal's z is -0.0 here and 0.0 in the test's literal
    assert np.array_equal(m.normals(), f32(fx["normals"]).reshape(-1, 3, 3))


@pytest.mark.parametrize("case", FIX["tri_loader"]["cases"], ids=lambda c: c["name"])
def test_tri_loader_cases(case):
    got = M.load_tri_text(case["input"])
    assert got.tobytes() == f32(case["expected"]).tobytes()


def test_unity_tri_asset():
    """tri_loader/tests/test_lib.rs: 12,583 triangles, all finite."""
    t = M.load_tri(os.path.join(GOLDEN, "unity.tri.gz"))
    assert t.shape == (FIX["unity_tri_count"]["expected"], 9)
    assert np.isfinite(t).all()


def test_tri_loader_errors():
    with pytest.raises(M.TriLoaderError) as e:
        M.load_tri_text("1 2 3 4 5 6 7 8\n")
    assert e.value.kind == "ExpectedFloat"  # the newline token where the ninth number belongs
    with pytest.raises(M.TriLoaderError) as e:
        M.load_tri_text("1 2 3 4 5 6 7 8")
    assert e.value.kind == "EndOfFile"
    with pytest.raises(M.TriLoaderError) as e:
        M.load_tri_text("1 2 3 4 5 6 7 8 x9\n")
    assert e.value.kind == "ExpectedFloat"
    with pytest.raises(M.TriLoaderError):
        M.load_tri_text("1 2 3 4 5 6 7 8 9\r\n")  # '\r' is a token of its own, not skipped
    with pytest.raises(M.TriLoaderError):
        M.load_tri_text("1 2 3 4 5 6 7 8 1_0\n")
    ok = M.load_tri_text("# comment\n\n 1 2 3 4 5 6 7 8 9 # trailing\n\t\\ 9 8 7 6 5 4 3 2 1\n")
    assert ok.tolist() == [[1, 2, 3, 4, 5, 6, 7, 8, 9], [9, 8, 7, 6, 5, 4, 3, 2, 1]]
    assert M.load_tri_text("").shape == (0, 9)


def test_obj_mesh_decoder():
    """test_obj_mesh.rs: the v/vt/vn cube, 12 triangles with their texture
    coordinates and normals, bit for bit."""
    fx = FIX["obj_mesh"]
    m = M.ObjMeshDecoder(io.StringIO(fx["input"])).read_mesh()
    assert m.len() == 12
    assert m.primitives().tobytes() == f32(fx["primitives"]).tobytes()
    assert m.tex_coords().tobytes() == f32(fx["tex_coords"]).tobytes()
    assert m.normals().tobytes() == f32(fx["normals"]).tobytes()


def test_cube_obj_asset_matches_cube_fixture():
    """bvhtracer/assets/cube.obj (v//vn corners) decodes to the vertices and
    faces of the cube fixture that pins the faithful build."""
    m = M.open_mesh(os.path.join(GOLDEN, "cube.obj"))
    cube = json.load(open(os.path.join(GOLDEN, "cube_obj.json")))
    V = np.array(cube["vertices"], np.float32)
    want = V[np.array(cube["faces"]) - 1].reshape(-1, 9)
    assert m.primitives().tobytes() == want.tobytes()
    assert (m.tex_coords() == 0).all() and set(np.abs(m.normals()).sum(2).ravel()) == {1.0}


def test_obj_polygons_and_relative_indices():
    doc = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf -4 -3 -2 -1\n"
    m = M.ObjMeshDecoder(io.StringIO(doc)).read_mesh()
    assert m.primitives().reshape(-1, 3, 3)[:, :, :2].tolist() == [[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]]
    with pytest.raises(M.ObjError):
        M.ObjMeshDecoder(io.StringIO("v 0 0 0\nf 1 2 3\n")).read_mesh()


def test_parse_f32_is_correctly_rounded():
    """Decimals within 2^-60 of the midpoint between 1 and 1 + 2^-23 round
    to binary64 exactly onto it; the binary32 result must still follow the
    exact decimal value."""
    mid = Fraction(1) + Fraction(1, 2**24)
    above = mid + Fraction(1, 2**60)
    below = mid - Fraction(1, 2**60)

    def dec(fr):  # exact decimal expansion of a dyadic fraction
        n, d = fr.numerator, fr.denominator
        k = d.bit_length() - 1
        digits = n * 5**k
        s = str(digits).rjust(k + 1, "0")
        return s[:-k] + "." + s[-k:]

    got = M.parse_f32([dec(above), dec(below), dec(mid), "0.1", "-2.5e-3", "inf", "nan"])
    assert float(got[0]) == 1 + 2**-23 and float(got[1]) == 1.0 and float(got[2]) == 1.0
    assert got[3] == np.float32(0.1) and got[4] == np.float32(-2.5e-3)
    assert np.isinf(got[5]) and np.isnan(got[6])
    assert float(np.float64(float(dec(above)))) == float(mid)  # the binary64 intermediate is the midpoint


def test_parse_f32_overflow_threshold():
    """Rust's parse rounds correctly at the overflow threshold FLT_MAX + 2^103
    (the midpoint between FLT_MAX and 2^128): below it FLT_MAX, at it or
    above infinity (the tie goes to the even significand, 2^128)."""
    from fractions import Fraction

    fmax = np.finfo(np.float32).max
    mid = Fraction(float(fmax)) + Fraction(2) ** 103
    assert str(mid) == "340282356779733661637539395458142568448"
    below = "3.4028235677973366e38"  # within a binary64 rounding of the midpoint
    assert Fraction(below) < mid and float(below) == float(mid)
    r = M.parse_f32([below, "-" + below, str(mid), "3.4028235677973367e38", "1e39"])
    assert r[0] == fmax and r[1] == -fmax
    assert np.isinf(r[2]) and np.isinf(r[3]) and np.isinf(r[4])
