"""ctypes binding of the C ABI in include/bvht.h (libbvht_hip.so).

The library is loaded from the package directory only; if it is missing the
import of any GPU entry point raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbvht_hip.so")

OK, EINVAL, ENOMEM, EDEVICE, ESTACK = 0, -1, -2, -3, -4
MISS = 0xFFFFFFFF
ORDER_ROWS, ORDER_TILES = 0, 1
BUILD_REFERENCE, BUILD_SAH = 0, 1
SGD_ALONG_NORMAL = 1
ACC_DEPTH, ACC_INTERSECTION, ACC_UV, ACC_NORMAL, ACC_TEXTURE = 0, 1, 2, 3, 4
SHADE_RADIANCE_TO_RGB, SHADE_INTERSECTION, SHADE_DEPTH_MAPPING = 0, 1, 2
ACCUM_PARAMS_DTYPE = np.dtype([("hit_value", "<f4", 3), ("miss_value", "<f4", 3)])
SHADER_PARAMS_DTYPE = np.dtype([("scale", "<f4"), ("offset", "<f4"), ("hit_rgba", "u1", 4), ("miss_rgba", "u1", 4)])

NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("left_first", "<u4"), ("bmax", "<f4", 3), ("count", "<u4")])
TNODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("left_right", "<u4"), ("bmax", "<f4", 3), ("blas", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("pad", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("packed", "<u4")])
INSTANCE_DTYPE = np.dtype([("m", "<f4", 16), ("minv", "<f4", 16), ("blas", "<u4"), ("pad", "<u4", 3)])
CAMERA_DTYPE = np.dtype([("tl", "<f4", 3), ("tr", "<f4", 3), ("bl", "<f4", 3), ("view_inv", "<f4", 16)])
STATS_DTYPE = np.dtype([("pairs", "<u4"), ("tests", "<u4"), ("tlas_pairs", "<u4"), ("inst_visits", "<u4")])
BODY_DTYPE = np.dtype([
    ("inverse_mass", "<f4"), ("inverse_inertia_tensor", "<f4", 9), ("linear_drag", "<f4"), ("angular_drag", "<f4"),
    ("position", "<f4", 3), ("orientation", "<f4", 4), ("velocity", "<f4", 3), ("rotation", "<f4", 3),
    ("inverse_inertia_tensor_world", "<f4", 9), ("motion", "<f4"), ("is_awake", "<u4"), ("can_sleep", "<u4"),
    ("transform", "<f4", 16), ("force_accumulator", "<f4", 3), ("torque_accumulator", "<f4", 3),
    ("acceleration", "<f4", 3), ("last_frame_acceleration", "<f4", 3), ("pad", "<f4", 3)])
FORCE_DTYPE = np.dtype([("kind", "<u4"), ("body", "<u4"), ("other", "<u4"), ("pad", "<u4"), ("v0", "<f4", 3),
                        ("v1", "<f4", 3), ("spring_constant", "<f4"), ("rest_length", "<f4")])
FORCE_NULL, FORCE_GRAVITY, FORCE_SPRING = 0, 1, 2

assert NODE_DTYPE.itemsize == 32 and TNODE_DTYPE.itemsize == 32 and RAY_DTYPE.itemsize == 32
assert HIT_DTYPE.itemsize == 24 and INSTANCE_DTYPE.itemsize == 144 and CAMERA_DTYPE.itemsize == 100
assert BODY_DTYPE.itemsize == 272 and FORCE_DTYPE.itemsize == 48

P = ctypes.c_void_p
PP = ctypes.POINTER(ctypes.c_void_p)
I32, U32, U64, F32, SZ = ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_size_t

# name -> (restype, argtypes); every symbol declared in include/bvht.h
SIGNATURES = {
    "bvht_create": (I32, [I32, PP]),
    "bvht_destroy": (I32, [P]),
    "bvht_last_error": (ctypes.c_char_p, []),
    "bvht_sync": (I32, [P]),
    "bvht_stream": (P, [P]),
    "bvht_device_alloc": (I32, [P, SZ, PP]),
    "bvht_device_free": (I32, [P, P]),
    "bvht_copy_to_device": (I32, [P, P, P, SZ]),
    "bvht_copy_to_host": (I32, [P, P, P, SZ]),
    "bvht_event_create": (I32, [P, PP]),
    "bvht_event_record": (I32, [P, P]),
    "bvht_event_elapsed": (I32, [P, P, ctypes.POINTER(F32)]),
    "bvht_event_destroy": (I32, [P]),
    "bvht_stream_copy": (I32, [P, P, P, SZ]),
    "bvht_blas_build": (I32, [P, P, U32, I32, U32, PP]),
    "bvht_blas_upload": (I32, [P, P, U32, U32, P, U32, PP]),
    "bvht_blas_download": (I32, [P, P, P, P]),
    "bvht_blas_info": (I32, [P, P, P, P, P]),
    "bvht_blas_set_vertices": (I32, [P, P, P, U32]),
    "bvht_blas_set_vertices_device": (I32, [P, P, P, U32]),
    "bvht_blas_permutation": (I32, [P, P, P]),
    "bvht_blas_refit": (I32, [P, P]),
    "bvht_blas_bounds": (I32, [P, P, P]),
    "bvht_blas_destroy": (I32, [P]),
    "bvht_intersect_pairs": (I32, [P, P, P, U32, P]),
    "bvht_trace_blas": (I32, [P, P, P, U32, P, P]),
    "bvht_trace_blas_count": (I32, [P, P, P, P, U32, P, P]),
    "bvht_set_trace_config": (I32, [P, I32, I32]),
    "bvht_get_trace_config": (I32, [P, ctypes.POINTER(I32), ctypes.POINTER(I32)]),
    "bvht_set_stack_limit": (I32, [P, U32]),
    "bvht_set_wave_schedule": (I32, [P, I32]),
    "bvht_raygen": (I32, [P, P, U32, U32, I32, P]),
    "bvht_bounce": (I32, [P, P, P, P, U32, U64, P, P]),
    "bvht_bounce_sorted": (I32, [P, P, P, P, U32, U64, P, P, P]),
    "bvht_depth": (I32, [P, P, U32, P]),
    "bvht_accumulate": (I32, [P, P, U32, U32, I32, I32, P, P]),
    "bvht_shade": (I32, [P, P, U32, I32, P, P]),
    "bvht_blas_set_attributes": (I32, [P, P, P, P, P, U32, U32]),
    "bvht_accumulate_blas": (I32, [P, P, P, U32, U32, I32, I32, P]),
    "bvht_accumulate_tlas": (I32, [P, P, P, U32, U32, I32, I32, P]),
    "bvht_tlas_create": (I32, [P, P, U32, P, U32, PP]),
    "bvht_tlas_set_objects": (I32, [P, P, P, U32, U32]),
    "bvht_tlas_refresh_bounds": (I32, [P, P]),
    "bvht_tlas_rebuild": (I32, [P, P]),
    "bvht_tlas_rebuild_async": (I32, [P, P]),
    "bvht_tlas_download": (I32, [P, P, P, P]),
    "bvht_tlas_destroy": (I32, [P]),
    "bvht_trace_tlas": (I32, [P, P, P, U32, P, P]),
    "bvht_trace_tlas_count": (I32, [P, P, P, P, U32, P, P]),
    "bvht_bounce_tlas": (I32, [P, P, P, P, U32, U64, P, P]),
    "bvht_bounce_tlas_sorted": (I32, [P, P, P, P, U32, U64, P, P, P]),
    "bvht_batch_create": (I32, [P, P, U32, PP]),
    "bvht_batch_destroy": (I32, [P]),
    "bvht_batch_refit": (I32, [P, P]),
    "bvht_batch_rebuild": (I32, [P, P]),
    "bvht_batch_set_rays": (I32, [P, P, P, U32]),
    "bvht_batch_config": (I32, [P, P, I32, I32, I32]),
    "bvht_batch_trace": (I32, [P, P, P, P]),
    "bvht_deform_wave": (I32, [P, P, P, F32, F32, F32, P]),
    "bvht_counter_add": (I32, [P, P, U32]),
    "bvht_mesh_create": (I32, [P, P, P, U32, P, U32, P, PP]),
    "bvht_mesh_destroy": (I32, [P]),
    "bvht_mesh_buffers": (I32, [P, PP, PP]),
    "bvht_mesh_update": (I32, [P, P, I32]),
    "bvht_mesh_sgd": (I32, [P, P, P, F32, F32, I32]),
    "bvht_trace_forward": (I32, [P, P, P, P, P, U32, P, P]),
    "bvht_trace_adjoint": (I32, [P, P, P, P, P, U32, P, P, F32, F32, I32, P, P, P]),
    "bvht_trace_camera_bounce": (I32, [P, P, P, U32, U32, I32, U64, P, P, P, P]),
    "bvht_trace_camera_bounce_noalloc": (I32, [P, P, P, U32, U32, I32, U64, P, P, P, P]),
    "bvht_world_create": (I32, [P, P, U32, P, U32, P]),
    "bvht_world_destroy": (I32, [P, P]),
    "bvht_world_set_bodies": (I32, [P, P, P, U32, U32]),
    "bvht_world_get_bodies": (I32, [P, P, P, U32, U32]),
    "bvht_world_start_frame": (I32, [P, P]),
    "bvht_world_run_physics": (I32, [P, P, F32]),
    "bvht_tlas_bind_world": (I32, [P, P, P, P, P]),
    "bvht_tlas_get_objects": (I32, [P, P, P, U32, U32]),
    "bvht_scene_run_physics": (I32, [P, P, F32]),
    "bvht_graph_begin": (I32, [P]),
    "bvht_graph_end": (I32, [P, PP]),
    "bvht_graph_launch": (I32, [P, P]),
    "bvht_graph_destroy": (I32, [P]),
}


class BvhtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bvht error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libbvht_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m bvhtracer_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().bvht_last_error()
        raise BvhtError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)
