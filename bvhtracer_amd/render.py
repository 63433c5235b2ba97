"""Host mirror of the reference's renderer (bvhtracer/src/renderer.rs) with
the whole frame on the GPU: camera rays in 8x8-tile order (bvht_raygen),
trace, accumulator into the device accumulation buffer at each ray's pixel
(bvht_accumulate), pixel shader into the device RGBA8 frame buffer
(bvht_shade).  Names and argument meaning follow renderer.rs.  The
normal-mapping and texture accumulators read the models' attributes (set by
ModelBuilder from a decoded Mesh and a texture).
"""
import numpy as np

from . import _lib as L
from .api import HIT_DTYPE, ORDER_TILES, RAY_DTYPE, default_context, raygen


class DepthAccumulator:
    """renderer.rs:176-194."""
    kind = L.ACC_DEPTH

    def params(self):
        return None


class IntersectionAccumulator:
    """renderer.rs:134-152."""
    kind = L.ACC_INTERSECTION

    def __init__(self, hit_value, miss_value):
        self.hit_value, self.miss_value = hit_value, miss_value

    def params(self):
        p = np.zeros(1, L.ACCUM_PARAMS_DTYPE)
        p["hit_value"][0] = self.hit_value
        p["miss_value"][0] = self.miss_value
        return p


class UvMappingAccumulator:
    """renderer.rs:225-244."""
    kind = L.ACC_UV

    def params(self):
        return None


class NormalMappingAccumulator:
    """renderer.rs:248-290 (needs the models' normals)."""
    kind = L.ACC_NORMAL

    def params(self):
        return None


class TextureMaterialAccumulator:
    """renderer.rs:293-334 (needs the models' texture coordinates and texture)."""
    kind = L.ACC_TEXTURE

    def params(self):
        return None


class RadianceToRgbShader:
    """renderer.rs:116-132."""
    kind = L.SHADE_RADIANCE_TO_RGB

    def params(self):
        return None


class IntersectionShader:
    """renderer.rs:155-173 (colours as RGBA bytes)."""
    kind = L.SHADE_INTERSECTION

    def __init__(self, hit_value, miss_value):
        self.hit_value, self.miss_value = hit_value, miss_value

    def params(self):
        p = np.zeros(1, L.SHADER_PARAMS_DTYPE)
        p["hit_rgba"][0] = self.hit_value
        p["miss_rgba"][0] = self.miss_value
        return p


class DepthMappingShader:
    """renderer.rs:196-222."""
    kind = L.SHADE_DEPTH_MAPPING

    def __init__(self, scale, offset):
        self.scale, self.offset = scale, offset

    def params(self):
        p = np.zeros(1, L.SHADER_PARAMS_DTYPE)
        p["scale"], p["offset"] = self.scale, self.offset
        return p


class RendererState:
    """RendererState (renderer.rs:77-101): accumulator, pixel shader, and the
    device accumulation buffer (float[3 W H]) and frame buffer (RGBA8,
    row-major), plus the frame's ray and hit buffers."""

    def __init__(self, accumulator, pixel_shader, width, height, ctx=None):
        if width % 8 or height % 8:
            raise ValueError("the tiled frame needs width and height multiples of 8")
        self.ctx = ctx or default_context()
        self.accumulator, self.pixel_shader = accumulator, pixel_shader
        self.width, self.height = width, height
        n = width * height
        self.d_rays = self.ctx.alloc(n * RAY_DTYPE.itemsize)
        self.d_hits = self.ctx.alloc(n * HIT_DTYPE.itemsize)
        self.d_accum = self.ctx.alloc(n * 12)
        self.d_frame = self.ctx.alloc(n * 4)
        self.d_frame.upload(np.tile(np.array([0, 0, 0, 255], np.uint8), n))  # FrameBuffer::from_fill(.., black)

    def accumulation_buffer(self):
        return self.d_accum.download(np.float32).reshape(self.height, self.width, 3)

    def frame_buffer(self):
        return self.d_frame.download(np.uint8).reshape(self.height, self.width, 4)


class PathTracer:
    """PathTracer::evaluate (renderer.rs:345-384): one camera ray per pixel
    in 8x8 tiles, u = x/W, v = y/H; the accumulator result at each pixel;
    then the pixel shader over the frame.  ``scene`` is anything with
    intersect_device (Scene, Tlas or Bvh); ``camera`` defaults to
    scene.active_camera().  Returns the number of rays traced."""

    def evaluate(self, state, scene, camera=None):
        ctx = state.ctx
        cam = camera if camera is not None else scene.active_camera()
        W, H = state.width, state.height
        n = W * H
        raygen(ctx, cam, W, H, ORDER_TILES, d_rays=state.d_rays)
        scene.intersect_device(state.d_rays, n, state.d_hits)
        kind = state.accumulator.kind
        if kind in (L.ACC_NORMAL, L.ACC_TEXTURE):
            tlas = getattr(scene, "tlas", scene if hasattr(scene, "rebuild_async") else None)
            if tlas is not None:
                L.check(L.lib().bvht_accumulate_tlas(ctx.handle, tlas.handle, state.d_hits.ptr, W, H, ORDER_TILES,
                                                     kind, state.d_accum.ptr))
            else:
                bvh = getattr(scene, "bvh", scene)
                L.check(L.lib().bvht_accumulate_blas(ctx.handle, bvh.handle, state.d_hits.ptr, W, H, ORDER_TILES,
                                                     kind, state.d_accum.ptr))
        else:
            ap = state.accumulator.params()
            L.check(L.lib().bvht_accumulate(ctx.handle, state.d_hits.ptr, W, H, ORDER_TILES, kind,
                                            L.ptr(ap) if ap is not None else None, state.d_accum.ptr))
        sp = state.pixel_shader.params()
        L.check(L.lib().bvht_shade(ctx.handle, state.d_accum.ptr, n, state.pixel_shader.kind,
                                   L.ptr(sp) if sp is not None else None, state.d_frame.ptr))
        return n


class Renderer:
    """Renderer (renderer.rs:386-398)."""

    def __init__(self, integrator):
        self.integrator = integrator

    def render(self, state, scene, camera=None):
        return self.integrator.evaluate(state, scene, camera)
