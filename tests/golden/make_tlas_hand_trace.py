"""Writes tests/golden/tlas_hand_trace.json: Tlas::rebuild
(bvhtracer/src/scene/tlas.rs:179-250) and Tlas::intersect (:123-177) traced
by hand on small scenes, with LeftRightIndex (:18-32) decoded as the
reference does (left() = high 16 bits = the SECOND argument of new()).

Every object is the unit cube [0,1]^3 translated by (x, 0, 0), so its world
box is [x, x+1] x [0,1] x [0,1].  The union of two such boxes spans
[min l, max r] in x and [0,1] in y and z; with extent X = max r - min l the
reference's area x*y + y*z + z*x is X + 1 + X = 2X + 1, an exact integer.
find_best_match scans b = 0..n-1, b != a, and keeps the first strict
minimum.  Notation: list = node_indices, cnt = node_index_count, area(i, j)
= area of the union of list[i]'s and list[j]'s boxes.  The expected node
arrays below follow from the steps written out in each case; nothing here is
computed by the oracle or the library.

Run from anywhere; it only writes the JSON beside this script.
"""
import json
import os
import sys

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tlas_hand_trace.json")


def lr(a, b):
    """LeftRightIndex::new(a, b).data = a + (b << 16) (tlas.rs:19-23)."""
    return a + (b << 16)


def leaf(x, blas):
    return {"bmin": [x, 0.0, 0.0], "bmax": [x + 1.0, 1.0, 1.0], "left_right": 0, "blas": blas}


def inner(x0, x1, left_right):
    # interior nodes keep TlasNode::default()'s blas = 0 (tlas.rs:48-56, 228-235)
    return {"bmin": [x0, 0.0, 0.0], "bmax": [x1, 1.0, 1.0], "left_right": left_right, "blas": 0}


FIVE = {
    "name": "five_objects_walk_replace_tie",
    "offsets_x": [0.0, 4.0, 6.0, 10.0, 12.0],
    "exercises": [
        "non-mutual best match: a = b; b = c (step 2, step 7)",
        "last-entry replacement list[b] = list[cnt-1] (step 3)",
        "equal-area tie, first index wins (step 4: slots 0 and 3 both 15)",
        "merge order new(list[a], list[b]) with a > b (node 7)",
        "root = copy of the last merge (node 0 = node 9)",
    ],
    "steps": [
        "leaves: list = [1,2,3,4,5], nodes 1..5 = objects 0..4 (x = 0,4,6,10,12), used = 6, cnt = 5",
        "1. a = 0; b = best(0): area(0,1)=2*5+1=11, (0,2)=15, (0,3)=23, (0,4)=27 -> b = 1",
        "2. c = best(1): (1,0)=11, (1,2)=[4,7]:7, (1,3)=15, (1,4)=19 -> c = 2 != a: a = 1, b = 2",
        "3. c = best(2): (2,0)=15, (2,1)=7, (2,3)=[6,11]:11, (2,4)=15 -> c = 1 == a: merge new(list[1]=2, list[2]=3) "
        "-> node 6 = [4,7], left_right = 2 + (3<<16); list[1] = 6; list[2] = list[4] = 5; cnt = 4; list = [1,6,5,4]",
        "4. b = best(1) (node 6 = [4,7]): (1,0)=[0,7]:15, (1,2)=[4,13]:19, (1,3)=[4,11]:15 -> tie 15, first index -> b = 0",
        "5. c = best(0) (node 1 = [0,1]): (0,1)=15, (0,2)=27, (0,3)=23 -> c = 1 == a: merge new(list[1]=6, list[0]=1) "
        "-> node 7 = [0,7], left_right = 6 + (1<<16); list[1] = 7; list[0] = list[3] = 4; cnt = 3; list = [4,7,5]",
        "6. b = best(1) (node 7 = [0,7]): (1,0)=[0,11]:23, (1,2)=[0,13]:27 -> b = 0",
        "7. c = best(0) (node 4 = [10,11]): (0,1)=23, (0,2)=[10,13]:7 -> c = 2 != a: a = 0, b = 2",
        "8. c = best(2) (node 5 = [12,13]): (2,0)=7, (2,1)=27 -> c = 0 == a: merge new(list[0]=4, list[2]=5) "
        "-> node 8 = [10,13], left_right = 4 + (5<<16); list[0] = 8; list[2] = list[2]; cnt = 2; list = [8,7]",
        "9. b = best(0): (0,1)=[0,13]:27 -> b = 1; c = best(1) = 0 == a: merge new(list[0]=8, list[1]=7) "
        "-> node 9 = [0,13], left_right = 8 + (7<<16); list[0] = 9; cnt = 1",
        "10. nodes[0] = nodes[list[a=0]] = node 9; nodes_used = 10",
        "(a last-index tie rule would take b = 3 at step 4 and merge nodes 4 and 5 into node 7 instead)",
    ],
    "nodes_used": 10,
    "nodes": [
        inner(0.0, 13.0, lr(8, 7)),
        leaf(0.0, 0), leaf(4.0, 1), leaf(6.0, 2), leaf(10.0, 3), leaf(12.0, 4),
        inner(4.0, 7.0, lr(2, 3)),
        inner(0.0, 7.0, lr(6, 1)),
        inner(10.0, 13.0, lr(4, 5)),
        inner(0.0, 13.0, lr(8, 7)),
    ],
}

REPLACE = {
    "name": "five_objects_replacement_orders_a_tie",
    "offsets_x": [10.0, 11.0, 13.0, 30.0, 8.0],
    "exercises": [
        "list[b] = list[cnt-1] moves object 4 ahead of object 2, and the next best match is a tie "
        "between them, so the first-index rule picks object 4 (step 3)",
    ],
    "steps": [
        "leaves: list = [1,2,3,4,5], nodes 1..5 = objects 0..4 (x = 10,11,13,30,8), used = 6, cnt = 5",
        "1. a = 0; b = best(0) (node 1 = [10,11]): (0,1)=[10,12]:5, (0,2)=[10,14]:9, (0,3)=[10,31]:43, "
        "(0,4)=[8,11]:7 -> b = 1",
        "2. c = best(1) (node 2 = [11,12]): (1,0)=5, (1,2)=[11,14]:7, (1,3)=[11,31]:41, (1,4)=[8,12]:9 -> c = 0 == a: "
        "merge new(1, 2) -> node 6 = [10,12], left_right = 1 + (2<<16); list[0] = 6; list[1] = list[4] = 5; "
        "cnt = 4; list = [6,5,3,4]",
        "3. b = best(0) (node 6 = [10,12]): (0,1)=[8,12]:9, (0,2)=[10,14]:9, (0,3)=[10,31]:43 -> tie 9, "
        "first index -> b = 1 (object 4)",
        "4. c = best(1) (node 5 = [8,9]): (1,0)=9, (1,2)=[8,14]:13, (1,3)=[8,31]:47 -> c = 0 == a: "
        "merge new(6, 5) -> node 7 = [8,12], left_right = 6 + (5<<16); list[0] = 7; list[1] = list[3] = 4; "
        "cnt = 3; list = [7,4,3]",
        "5. b = best(0) (node 7 = [8,12]): (0,1)=[8,31]:47, (0,2)=[8,14]:13 -> b = 2",
        "6. c = best(2) (node 3 = [13,14]): (2,0)=13, (2,1)=[13,31]:37 -> c = 0 == a: merge new(7, 3) "
        "-> node 8 = [8,14], left_right = 7 + (3<<16); list[0] = 8; list[2] = list[2]; cnt = 2; list = [8,4]",
        "7. b = best(0) = 1; c = best(1) = 0 == a: merge new(8, 4) -> node 9 = [8,31], left_right = 8 + (4<<16); cnt = 1",
        "8. nodes[0] = node 9; nodes_used = 10",
        "(removing entry b by shifting the list would leave [6,3,4,5]; step 3 would then pick object 2 "
        "and node 7 would be new(6, 3))",
    ],
    "nodes_used": 10,
    "nodes": [
        inner(8.0, 31.0, lr(8, 4)),
        leaf(10.0, 0), leaf(11.0, 1), leaf(13.0, 2), leaf(30.0, 3), leaf(8.0, 4),
        inner(10.0, 12.0, lr(1, 2)),
        inner(8.0, 12.0, lr(6, 5)),
        inner(8.0, 14.0, lr(7, 3)),
        inner(8.0, 31.0, lr(8, 4)),
    ],
}

THREE = {
    "name": "three_objects_a_beyond_count",
    "offsets_x": [0.0, 5.0, 3.0],
    "exercises": [
        "the first merge has a = cnt-1, so list[b] = list[cnt-1] copies the new node into slot b "
        "and a is left at the old last slot, outside 0..cnt (step 3)",
        "best(a) with a outside 0..cnt scans every slot, matching the new node with its own copy (step 4)",
        "merge with a > b: new(list[2], list[1]) (node 4)",
    ],
    "steps": [
        "leaves: list = [1,2,3], nodes 1..3 = objects 0..2 (x = 0,5,3), used = 4, cnt = 3",
        "1. a = 0; b = best(0): (0,1)=[0,6]:13, (0,2)=[0,4]:9 -> b = 2",
        "2. c = best(2) (node 3 = [3,4]): (2,0)=9, (2,1)=[3,6]:7 -> c = 1 != a: a = 2, b = 1",
        "3. c = best(1) (node 2 = [5,6]): (1,0)=13, (1,2)=7 -> c = 2 == a: merge new(list[2]=3, list[1]=2) "
        "-> node 4 = [3,6], left_right = 3 + (2<<16); list[2] = 4; list[1] = list[2] = 4; cnt = 2; list = [1,4,(4)]",
        "4. b = best(list, 2, a=2): slot 0: node 4 with node 1 = [0,6]:13; slot 1: node 4 with itself = [3,6]:7 -> b = 1",
        "5. c = best(1) (node 4): (1,0)=13 -> c = 0 != a=2: a = 1, b = 0",
        "6. c = best(0) (node 1): (0,1)=13 -> c = 1 == a: merge new(list[1]=4, list[0]=1) "
        "-> node 5 = [0,6], left_right = 4 + (1<<16); list[1] = 5; list[0] = list[1] = 5; cnt = 1",
        "7. nodes[0] = nodes[list[a=1]] = node 5; nodes_used = 6",
    ],
    "nodes_used": 6,
    "nodes": [
        inner(0.0, 6.0, lr(4, 1)),
        leaf(0.0, 0), leaf(5.0, 1), leaf(3.0, 2),
        inner(3.0, 6.0, lr(3, 2)),
        inner(0.0, 6.0, lr(4, 1)),
    ],
}

# The ray of the identical-object cases and its hit on the model triangle
# (0,0,0), (1,0,0), (0,1,1) by Triangle::intersect (triangle.rs:41-72):
# e1 = (1,0,0), e2 = (0,1,1), h = d x e2 = (-1,0,0), a = e1.h = -1, f = -1,
# s = o - v0 = (0.25,0.25,-5), u = f (s.h) = 0.25, q = s x e1 = (0,-5,-0.25),
# v = f (d.q) = 0.25, t = f (e2.q) = 5.25.  Every value is exact in binary32.
TRIANGLE = [0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0]
RAY = {"o": [0.25, 0.25, -5.0], "d": [0.0, 0.0, 1.0]}
HIT = {"t": 5.25, "u": 0.25, "v": 0.25, "prim": 0}

TWO_SAME = {
    "name": "two_identical_objects",
    "offsets_x": [0.0, 0.0],
    "exercises": [
        "LeftRightIndex::new(1, 2): left() = 2 (object 1), right() = 1 (object 0)",
        "Tlas::intersect: equal child distances go right (tlas.rs:149-155), and equal t is not accepted (:131)",
    ],
    "steps": [
        "leaves: list = [1,2], used = 3, cnt = 2; every area is 3 (unit cube)",
        "1. a = 0; b = best(0) = 1; c = best(1) = 0 == a: merge new(1, 2) -> node 3 = [0,1], left_right = 1 + (2<<16)",
        "2. nodes[0] = node 3; nodes_used = 4",
        "trace: root's left() = node 2 (object 1), right() = node 1 (object 0); both boxes give tmin = 5; "
        "5 < 5 is false, so the near child is right = object 0, and object 1 is pushed; object 0 hits at t = 5.25; "
        "object 1 hits at 5.25, not below 5.25, so the hit stays object 0",
    ],
    "nodes_used": 4,
    "nodes": [
        inner(0.0, 1.0, lr(1, 2)),
        leaf(0.0, 0), leaf(0.0, 1),
        inner(0.0, 1.0, lr(1, 2)),
    ],
    "hit_inst": 0,
    "stats": {"tlas_pairs": 1, "inst_visits": 2},
}

THREE_SAME = {
    "name": "three_identical_objects",
    "offsets_x": [0.0, 0.0, 0.0],
    "exercises": [
        "ties on every area: b = best(0) = 1 (first index), then the replacement list[1] = list[2]",
        "right() descends first on equal distances at both levels",
    ],
    "steps": [
        "leaves: list = [1,2,3], used = 4, cnt = 3; every area is 3",
        "1. a = 0; b = best(0): (0,1)=3, (0,2)=3 -> b = 1; c = best(1): (1,0)=3, (1,2)=3 -> c = 0 == a: "
        "merge new(1, 2) -> node 4, left_right = 1 + (2<<16); list[0] = 4; list[1] = list[2] = 3; cnt = 2",
        "2. b = best(0) = 1; c = best(1) = 0 == a: merge new(4, 3) -> node 5, left_right = 4 + (3<<16); cnt = 1",
        "3. nodes[0] = node 5; nodes_used = 6",
        "trace: root: left() = node 3 (object 2), right() = node 4 -> go to node 4, push node 3; "
        "node 4: left() = node 2 (object 1), right() = node 1 (object 0) -> object 0 first, push node 2; "
        "object 0 hits at 5.25; the pops of objects 1 and 2 hit at 5.25 and are not accepted",
    ],
    "nodes_used": 6,
    "nodes": [
        inner(0.0, 1.0, lr(4, 3)),
        leaf(0.0, 0), leaf(0.0, 1), leaf(0.0, 2),
        inner(0.0, 1.0, lr(1, 2)),
        inner(0.0, 1.0, lr(4, 3)),
    ],
    "hit_inst": 0,
    "stats": {"tlas_pairs": 2, "inst_visits": 3},
}


def main():
    fixture = {
        "source": "bvhtracer/src/scene/tlas.rs:18-32 (LeftRightIndex), :123-177 (intersect), "
                  ":179-250 (find_best_match, rebuild)",
        "model_triangle": TRIANGLE,
        "ray": RAY,
        "hit": HIT,
        "cases": [FIVE, REPLACE, THREE, TWO_SAME, THREE_SAME],
    }
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    sys.exit(main())
