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
