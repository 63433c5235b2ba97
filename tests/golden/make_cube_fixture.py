"""Writes tests/golden/cube_obj.json from the reference's asset
bvhtracer/assets/cube.obj (vertices and faces, data only) plus the expected
faithful-build result traced by hand from bvhtracer/src/model/bvh.rs:333-541:
final triangle order [T3,T2,T9,T8,T1,T0,T4,T10,T11,T5,T7,T6] (so T4, the +y
face hit by the first ray of bvhtracer/tests/test_scene_cube.rs, is primitive
6 and T5 is primitive 9, the ids that test asserts at :124-135 and :215-226)
and nodes 0 -> (2,3), 2 = {T3,T2}, 3 -> (4,5), 4 = {T9,T8}, 5 -> (6,7),
6 = {T1,T0}, 7 = the remaining six.

Run in a container that has the reference checkout at /root/reference.
"""
import json
import os
import sys

SRC = "/root/reference/bvhtracer/assets/cube.obj"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cube_obj.json")


def main():
    verts, faces = [], []
    with open(SRC) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                verts.append([float(x) for x in p[1:4]])
            elif p[0] == "f":
                faces.append([int(x.split("/")[0]) for x in p[1:4]])
    fixture = {
        "source": "bvhtracer/assets/cube.obj",
        "vertices": verts,
        "faces": faces,
        "expected_order": [3, 2, 9, 8, 1, 0, 4, 10, 11, 5, 7, 6],
        "expected_nodes_used": 8,
        "expected_left_count": [[2, 0], [0, 0], [0, 2], [4, 0], [2, 2], [6, 0], [4, 2], [6, 6]],
    }
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    sys.exit(main())
