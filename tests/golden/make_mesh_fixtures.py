"""Writes the mesh-ingestion fixtures of tests/golden/ from the reference
checkout at /root/reference (data only: the mesh assets its tests and
benchmarks load, and the input texts and expected values of its decoder
tests).

  unity.tri.gz, bigben.tri.gz  bvhtracer/assets/{unity,bigben}.tri, gzipped
                               (tri_loader/tests/test_lib.rs,
                               benches/bench_bvh_build.rs,
                               benches/bench_bvh_refit_rebuild.rs)
  cube.obj                     bvhtracer/assets/cube.obj
                               (tests/test_scene_cube.rs, examples/cube.rs)
  mesh_decoders.json           input documents and expected meshes of
                               tests/test_tri_mesh.rs, tests/test_obj_mesh.rs
                               and tri_loader/src/loader.rs's unit tests,
                               plus the element count of test_lib.rs

Expected values are taken as the numbers written in those tests (decimal
literals, to be rounded to f32 by the reader, as Rust does).
"""
import gzip
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?(?:\d+\.?\d*(?:e[-+]?\d+)?)(?:_f32)?"


def raw_string(src, fn_name):
    """The r"..." literal returned by fn_name (test input document)."""
    body = src[src.index(f"fn {fn_name}"):]
    start = body.index('r"') + 2
    return body[start:body.index('"', start)]


def vectors(block, ctor):
    """All ctor(a, b[, c]) literals in block, as lists of decimal strings."""
    out = []
    for m in re.finditer(re.escape(ctor) + r"\(([^()]*)\)", block):
        args = [a.strip().replace("_f32", "") for a in m.group(1).split(",") if a.strip()]
        out.append(args)
    return out


def expected_mesh(src):
    """primitives / tex_coords / normals of the MeshBuilder chain in fn mesh()."""
    body = src[src.index("fn mesh()"):]
    body = body[:body.index(".build()")]
    prims, texs, norms = [], [], []
    for part in body.split(".with_primitive(")[1:]:
        tri = part[part.index("Triangle::new("):part.index("TextureCoordinates")]
        tex = part[part.index("TextureCoordinates"):part.index("Normals")]
        nrm = part[part.index("Normals"):]
        prims.append(vectors(tri, "Vector3::new"))
        tv = vectors(tex, "Vector2::new")
        texs.append(tv if tv else [["0", "0"]] * 3)  # Vector2::zero()
        nv = vectors(nrm, "Vector3::new")
        norms.append(nv)
    return {"primitives": prims, "tex_coords": texs, "normals": norms}


def loader_cases(src):
    cases = []
    tests = src[src.index("mod tests"):]
    for m in re.finditer(r"fn (test_\w+)\(\)", tests):
        body = tests[m.end():]
        body = body[:body.find("#[test]") if "#[test]" in body else len(body)]
        data = body[body.index('r"') + 2:]
        data = data[:data.index('"')]
        tris = [vectors(t, "Vertex::new") for t in body.split("Triangle::new(")[1:]]
        cases.append({"name": m.group(1), "input": data, "expected": tris})
    return cases


def main():
    for name in ("unity.tri", "bigben.tri"):
        with open(os.path.join(REF, "bvhtracer/assets", name), "rb") as f, \
                gzip.open(os.path.join(HERE, name + ".gz"), "wb", compresslevel=9) as g:
            shutil.copyfileobj(f, g)
    shutil.copy(os.path.join(REF, "bvhtracer/assets/cube.obj"), os.path.join(HERE, "cube.obj"))
    tri_src = open(os.path.join(REF, "bvhtracer/tests/test_tri_mesh.rs")).read()
    obj_src = open(os.path.join(REF, "bvhtracer/tests/test_obj_mesh.rs")).read()
    loader_src = open(os.path.join(REF, "bvhtracer/tri_loader/src/loader.rs")).read()
    lib_src = open(os.path.join(REF, "bvhtracer/tri_loader/tests/test_lib.rs")).read()
    count = int(re.search(r"let expected = (\d+);", lib_src).group(1))
    fixture = {
        "tri_mesh": {"source": "bvhtracer/tests/test_tri_mesh.rs", "input": raw_string(tri_src, "test_case"),
                     **expected_mesh(tri_src)},
        "obj_mesh": {"source": "bvhtracer/tests/test_obj_mesh.rs", "input": raw_string(obj_src, "test_case"),
                     **expected_mesh(obj_src)},
        "tri_loader": {"source": "bvhtracer/tri_loader/src/loader.rs", "cases": loader_cases(loader_src)},
        "unity_tri_count": {"source": "bvhtracer/tri_loader/tests/test_lib.rs", "expected": count},
    }
    with open(os.path.join(HERE, "mesh_decoders.json"), "w") as f:
        json.dump(fixture, f, indent=1)


if __name__ == "__main__":
    main()
