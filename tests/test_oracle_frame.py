"""Pins of the frame-loop oracle (renderer.rs): the tile loop's pixel
addresses against a literal transcription of PathTracer::evaluate's loop,
and the pixel shaders against values computed with Python integers under
Rust's cast and wrapping rules."""
import numpy as np

from oracle import frame

F32_MAX = float(np.finfo(np.float32).max)


def test_tile_order_matches_the_reference_loop():
    """PathTracer::evaluate (renderer.rs:345-368) at 640x640 as written
    (tile_count_x = tile_count_y = 80, y = tile / tile_count_y)."""
    W = H = 640
    addr = []
    for tile in range(80 * 80):
        x, y = tile % 80, tile // 80
        for v in range(8):
            for u in range(8):
                addr.append((x * 8 + u) + (y * 8 + v) * W)
    assert np.array_equal(frame.tile_pixel_order(W, H), np.array(addr))


def depth_color(t, scale, offset):
    """DepthMappingShader::evaluate (renderer.rs:207-222) with Python ints:
    f32 arithmetic, saturating `as i32`, two's-complement `as u32`, wrapping
    u32 subtraction and multiplication."""
    f = np.float32((np.float32(t) - np.float32(offset)) * np.float32(scale))
    q = 0 if f != f else int(max(min(np.trunc(f), 2**31 - 1), -2**31))
    c = ((255 - (q % 2**32)) % 2**32) * 0x010101 % 2**32
    return (c >> 16) & 255, (c >> 8) & 255, c & 255


def test_depth_mapping_shader_closed_forms():
    ts = [3.0, 4.0, 3.5, 6.1875, 6.5, 2.0, -1e6, 1e30, F32_MAX, float("nan"), 5.99]
    acc = np.array([[t, t, t] for t in ts], np.float32)
    got = frame.shade(acc, "depth", scale=80.0, offset=3.0)
    for k, t in enumerate(ts):
        want = (0, 0, 0) if not (np.float32(t) < np.float32(F32_MAX)) else depth_color(t, 80.0, 3.0)
        assert tuple(got[k, :3]) == want, (t, got[k], want)
        assert got[k, 3] == 255
    assert tuple(got[0, :3]) == (255, 255, 255) and tuple(got[1, :3]) == (175, 175, 175)


def test_radiance_and_intersection_shaders():
    acc = np.array([[0.5, 1.2, -0.3], [np.nan, 1.0, 0.999], [0, 0, 0], [0, 0, 1e-30]], np.float32)
    rgb = frame.shade(acc, "radiance")
    assert rgb[0].tolist() == [127, 255, 0, 255]
    assert rgb[1].tolist() == [0, 255, 254, 255]
    hit, miss = (1, 2, 3, 4), (9, 8, 7, 6)
    sh = frame.shade(acc, "intersection", hit_rgba=hit, miss_rgba=miss)
    assert [tuple(r) for r in sh] == [hit, hit, miss, hit]


def test_accumulators_on_hand_made_hits():
    hits = np.zeros(64, [("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4"), ("inst", "<u4"),
                         ("packed", "<u4")])
    hits["prim"] = np.where(np.arange(64) % 3 == 0, 0xFFFFFFFF, 7)
    hits["t"], hits["u"], hits["v"] = np.arange(64), 0.25, 0.5
    d = frame.accumulate(hits, 8, 8, "depth")
    assert d[0, 0] == np.float32(F32_MAX) and d[1].tolist() == [1.0, 1.0, 1.0]
    uv = frame.accumulate(hits, 8, 8, "uv")
    assert uv[1].tolist() == [0.25, 0.5, 0.25] and uv[0].tolist() == [0, 0, 0]
    it = frame.accumulate(hits, 8, 8, "intersection", (1, 1, 1), (0, 0, 0.5))
    assert it[0].tolist() == [0, 0, 0.5] and it[1].tolist() == [1, 1, 1]


HIT = [("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("packed", "<u4")]


def test_normal_mapping_closed_form():
    """Corner normals all +y under M = 2 I (+ a translation, which vectors
    ignore): world normal (0, 2, 0) -> (0, 1, 0) -> (0.5, 1, 0.5) exactly;
    a miss gives 0."""
    hits = np.zeros(2, HIT)
    hits["prim"] = [0, 0xFFFFFFFF]
    hits["u"], hits["v"] = 0.25, 0.5
    model = {"normals": np.tile(np.array([0, 1, 0], np.float32), (1, 3, 1))}
    M = np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32)
    M[3, :3] = -1.0  # column 3 = translation in the column-major layout
    got = frame.accumulate_attributes(hits, 2, 1, "normal", [model], [(M.reshape(16), 0)], order="rows")
    assert got.tolist() == [[0.5, 1.0, 0.5], [0.0, 0.0, 0.0]]


def test_texture_lookup_closed_form():
    """Nearest texel at ((u W) as usize % W, (v H) as usize % H): corner
    coordinates equal, so the lookup sees them exactly; negative and NaN
    coordinates go to texel 0, coordinates past 1 wrap."""
    tex = np.arange(4 * 2 * 3, dtype=np.uint8).reshape(2, 4, 3)  # H = 2, W = 4
    uvs = [(0.3, 0.6), (-0.5, 0.2), (1.3, 2.75), (np.nan, 0.9), (0.999, 0.0)]
    want_px = [(1, 1), (0, 0), (1, 1), (0, 1), (3, 0)]  # (iu, iv)
    hits = np.zeros(len(uvs), HIT)
    hits["prim"] = np.arange(len(uvs))
    hits["u"], hits["v"] = 0.2, 0.3
    tcs = np.array([[uv, uv, uv] for uv in uvs], np.float32)
    got = frame.accumulate_attributes(hits, len(uvs), 1, "texture", [{"tex_coords": tcs, "texture": tex}],
                                      order="rows")
    for k, (iu, iv) in enumerate(want_px):
        assert got[k].tolist() == (tex[iv, iu].astype(np.float32) / 256).tolist(), (uvs[k], got[k])
