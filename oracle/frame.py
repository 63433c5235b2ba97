"""Frame-loop oracle (TEST INFRASTRUCTURE ONLY): numpy restatement of the
reference renderer's tiled loop, hit-record accumulators and pixel shaders
(lambdaxymox/bvhtracer, bvhtracer/src/renderer.rs), used as the checker of
the GPU frame path.  Integer and byte arithmetic follows Rust's semantics:
saturating float -> int casts and, for DepthMappingShader, wrapping u32
arithmetic (a release build; a debug build panics where it would wrap).
"""
import numpy as np

F32_MAX = np.float32(np.finfo(np.float32).max)
MISS = 0xFFFFFFFF


def tile_pixel_order(width, height):
    """Pixel address y*W + x of the j-th ray of PathTracer::evaluate's loop
    (renderer.rs:349-368): 8x8 tiles in row-major tile order, pixels in
    row-major order inside a tile."""
    tiles_x = width // 8
    j = np.arange(width * height)
    tile, w = j // 64, j % 64
    x = (tile % tiles_x) * 8 + (w % 8)
    y = (tile // tiles_x) * 8 + (w // 8)
    return y * width + x


def accumulate(hits, width, height, kind, hit_value=None, miss_value=None, order="tiles"):
    """Accumulator::evaluate for every ray, stored at its pixel address.
    kind: 'depth' (renderer.rs:183-194), 'intersection' (:144-152),
    'uv' (:228-243).  Returns float32 (H*W, 3)."""
    hits = np.asarray(hits)
    hit = hits["prim"] != MISS
    n = width * height
    val = np.zeros((n, 3), np.float32)
    if kind == "depth":
        t = np.where(hit, hits["t"], F32_MAX).astype(np.float32)
        val[:] = t[:, None]
    elif kind == "intersection":
        val[:] = np.where(hit[:, None], np.asarray(hit_value, np.float32)[None, :],
                          np.asarray(miss_value, np.float32)[None, :])
    elif kind == "uv":
        u, v = hits["u"].astype(np.float32), hits["v"].astype(np.float32)
        one = np.float32(1.0)
        val[:, 0] = np.where(hit, u, 0)
        val[:, 1] = np.where(hit, v, 0)
        val[:, 2] = np.where(hit, one - (u + v), 0)
    else:
        raise ValueError(kind)
    out = np.zeros((n, 3), np.float32)
    addr = tile_pixel_order(width, height) if order == "tiles" else np.arange(n)
    out[addr] = val
    return out


def _sat_cast(x, lo, hi):
    """Rust `f32 as int`: truncation toward zero, NaN -> 0, clamped."""
    x = np.asarray(x, np.float64)
    r = np.trunc(np.nan_to_num(x, nan=0.0, posinf=hi, neginf=lo))
    return np.clip(r, lo, hi).astype(np.int64)


def shade(accum, kind, scale=0.0, offset=0.0, hit_rgba=(0, 0, 0, 0), miss_rgba=(0, 0, 0, 0)):
    """PixelShader::evaluate per pixel: 'radiance' (renderer.rs:125-132),
    'intersection' (:166-173), 'depth' (:207-222).  Returns uint8 (n, 4)."""
    a = np.asarray(accum, np.float32).reshape(-1, 3)
    n = a.shape[0]
    out = np.zeros((n, 4), np.uint8)
    if kind == "radiance":
        for c in range(3):
            out[:, c] = np.minimum(255, _sat_cast(np.float32(255.0) * a[:, c], 0, 255))
        out[:, 3] = 255
    elif kind == "intersection":
        nonzero = (a != 0).any(axis=1)
        out[:] = np.where(nonzero[:, None], np.asarray(hit_rgba, np.uint8)[None, :],
                          np.asarray(miss_rgba, np.uint8)[None, :])
    elif kind == "depth":
        t = a[:, 0]
        with np.errstate(over="ignore", invalid="ignore"):
            f = (t - np.float32(offset)) * np.float32(scale)
        q = _sat_cast(f, -2**31, 2**31 - 1) & 0xFFFFFFFF
        c = ((255 - q) & 0xFFFFFFFF) * 0x010101 & 0xFFFFFFFF
        rgb = np.stack([(c & 0xFF0000) >> 16, (c & 0xFF00) >> 8, c & 0xFF], axis=1)
        out[:, :3] = np.where((t < F32_MAX)[:, None], rgb, 0).astype(np.uint8)
        out[:, 3] = 255
    else:
        raise ValueError(kind)
    return out


def _usize(x):
    """Rust `f32 as usize` (64-bit), then kept as Python ints."""
    x = np.asarray(x, np.float64)
    out = []
    for v in x.ravel():
        if not v > 0:
            out.append(0)
        elif v >= 2.0**64:
            out.append(2**64 - 1)
        else:
            out.append(int(v))
    return out


def accumulate_attributes(hits, width, height, kind, models, objects=None, order="tiles"):
    """NormalMappingAccumulator (renderer.rs:248-290) and
    TextureMaterialAccumulator (renderer.rs:293-334, TextureMaterial::evaluate
    materials/material.rs:44-52) per ray, stored at its pixel address.
    models: list of dicts with 'normals' (n, 3, 3), 'tex_coords' (n, 3, 2),
    'texture' (H, W, 3) uint8, triangles in build order.  objects: list of
    (M as 16 column-major floats, model index), or None for a single-BLAS
    trace (identity transform, model 0).  Indices are the hits' true prim and
    inst.  binary32 in the reference's operation order."""
    hits = np.asarray(hits)
    n = width * height
    val = np.zeros((n, 3), np.float32)
    one, half = np.float32(1.0), np.float32(0.5)
    ident = np.eye(4, dtype=np.float32).reshape(16)
    for j in range(n):
        h = hits[j]
        if h["prim"] == MISS:
            continue
        if objects is None:
            m16, mi = ident, 0
        else:
            if h["inst"] >= len(objects):
                continue
            m16, mi = objects[h["inst"]]
        mdl = models[mi]
        u, v = np.float32(h["u"]), np.float32(h["v"])
        w = (one - u) - v
        p = int(h["prim"])
        if kind == "normal":
            c = np.asarray(mdl["normals"], np.float32)[p]
            m = ((c[0] * w) + (c[1] * u)) + (c[2] * v)
            M = np.asarray(m16, np.float32)
            ws = np.array([((M[0 + r] * m[0] + M[4 + r] * m[1]) + M[8 + r] * m[2]) + M[12 + r] * np.float32(0)
                           for r in range(3)], np.float32)
            ln = np.sqrt((ws[0] * ws[0] + ws[1] * ws[1]) + ws[2] * ws[2])
            val[j] = (ws / ln + one) * half
        elif kind == "texture":
            c = np.asarray(mdl["tex_coords"], np.float32)[p]
            tex = np.asarray(mdl["texture"], np.uint8)
            th, tw = tex.shape[0], tex.shape[1]
            uv = ((c[0] * w) + (c[1] * u)) + (c[2] * v)
            iu = _usize(uv[0] * np.float32(tw))[0] % tw
            iv = _usize(uv[1] * np.float32(th))[0] % th
            s = np.float32(1.0) / np.float32(256.0)
            val[j] = tex[iv, iu].astype(np.float32) * s
        else:
            raise ValueError(kind)
    out = np.zeros((n, 3), np.float32)
    addr = tile_pixel_order(width, height) if order == "tiles" else np.arange(n)
    out[addr] = val
    return out
