"""bvhtracer_amd -- MI355X (gfx950) ray/BVH traversal hot path.

A drop-in for the traversal path of lambdaxymox/bvhtracer: hand-written HIP
kernels behind the C ABI in include/bvht.h (libbvht_hip.so), with a host
mirror of the reference's Bvh / BvhBuilder / Tlas / SceneObject / Ray API.
"""
from ._lib import (BUILD_REFERENCE, BUILD_SAH, CAMERA_DTYPE, HIT_DTYPE, INSTANCE_DTYPE, MISS, NODE_DTYPE,
                   ORDER_ROWS, ORDER_TILES, RAY_DTYPE, STATS_DTYPE, TNODE_DTYPE, BvhtError)
from .api import (Batch, Bvh, BvhBuilder, Context, DeviceArray, Graph, Mesh, Ray, SceneObject, Tlas, TlasBuilder,
                  bounce, bounce_sorted, bounce_tlas, bounce_tlas_sorted, counter_add, default_context, deform_wave, raygen, trace_adjoint, trace_camera_bounce, trace_forward,
                  transform_gradient)

from .render import (DepthAccumulator, DepthMappingShader, IntersectionAccumulator, IntersectionShader,
                     NormalMappingAccumulator, PathTracer, RadianceToRgbShader, Renderer, RendererState,
                     TextureMaterialAccumulator, UvMappingAccumulator)
from .scene import (ModelBuilder, ModelInstance, Scene, SceneBuilder, SceneObjectBuilder, Triangle,
                    intersect_pairs)
from .physics import Gravity, NullForce, RigidBody, RigidBodyInstance, Spring, World

__all__ = [
    "BUILD_REFERENCE", "BUILD_SAH", "CAMERA_DTYPE", "HIT_DTYPE", "INSTANCE_DTYPE", "MISS", "NODE_DTYPE",
    "ORDER_ROWS", "ORDER_TILES", "RAY_DTYPE", "STATS_DTYPE", "TNODE_DTYPE", "BvhtError",
    "Batch", "Bvh", "BvhBuilder", "Context", "DeviceArray", "Graph", "Mesh", "Ray", "SceneObject", "Tlas",
    "TlasBuilder", "bounce", "bounce_sorted", "bounce_tlas", "bounce_tlas_sorted", "counter_add", "default_context", "deform_wave", "raygen",
    "trace_adjoint", "trace_camera_bounce", "trace_forward", "transform_gradient",
    "DepthAccumulator", "DepthMappingShader", "IntersectionAccumulator", "IntersectionShader",
    "NormalMappingAccumulator", "PathTracer", "RadianceToRgbShader", "Renderer", "RendererState",
    "TextureMaterialAccumulator", "UvMappingAccumulator",
    "ModelBuilder", "ModelInstance", "Scene", "SceneBuilder", "SceneObjectBuilder", "Triangle", "intersect_pairs",
    "Gravity", "NullForce", "RigidBody", "RigidBodyInstance", "Spring", "World",
]
