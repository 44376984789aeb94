"""ctypes binding of ``liblucent_hip.so`` (the C ABI in include/lucent_hip.h).

The library must have been built (``python -m lucent_amd.build``); there is
no fallback of any kind — importing a compute entry point without the
library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblucent_hip.so"

LUCENT_OK = 0
LUCENT_QUERY_CLOSEST = 0
LUCENT_QUERY_ANY = 1
LUCENT_PREC_EXACT_FP64 = 0
LUCENT_PREC_GLSL_FP32 = 2
LUCENT_MISS = 0xFFFFFFFF

STATUS_NAMES = {
    -1: "LUCENT_E_INVALID",
    -2: "LUCENT_E_HIP",
    -3: "LUCENT_E_NO_DEVICE",
    -4: "LUCENT_E_OUT_OF_MEMORY",
    -5: "LUCENT_E_UNSUPPORTED",
    -6: "LUCENT_E_OVERFLOW",
}


class LucentError(RuntimeError):
    """A failed C-ABI call; carries the status code and lucent_last_error()."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class PackedSceneStruct(C.Structure):
    _fields_ = [
        ("tri_geometry", C.c_void_p),
        ("tri_attributes", C.c_void_p),
        ("tri_count", C.c_uint32),
        ("bvh_nodes", C.c_void_p),
        ("bvh_node_count", C.c_uint32),
        ("tri_indices", C.c_void_p),
        ("materials", C.c_void_p),
        ("material_count", C.c_uint32),
        ("spheres", C.c_void_p),
        ("sphere_materials", C.c_void_p),
        ("sphere_lights", C.c_void_p),
        ("sphere_count", C.c_uint32),
        ("lights", C.c_void_p),
        ("light_count", C.c_uint32),
        ("ground_material", C.c_int32),
        ("ground_height", C.c_float),
    ]


class SceneStatsStruct(C.Structure):
    _fields_ = [
        ("tri_count", C.c_uint32),
        ("node_count", C.c_uint32),
        ("node4_count", C.c_uint32),
        ("depth", C.c_uint32),
        ("device_bytes", C.c_uint64),
    ]


class SunStruct(C.Structure):
    _fields_ = [("elevation", C.c_double), ("azimuth", C.c_double), ("color", C.c_double * 3),
                ("intensity", C.c_double), ("angular_radius", C.c_double)]


class SkyStruct(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("zenith", C.c_double * 3), ("horizon", C.c_double * 3),
                ("ground", C.c_double * 3), ("intensity", C.c_double)]


class EnvironmentStruct(C.Structure):
    _fields_ = [("map_rgb", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("rotation", C.c_double),
                ("intensity", C.c_double), ("background", C.c_int32), ("backdrop", C.c_double * 3)]


class LightingStruct(C.Structure):
    _fields_ = [("sun", SunStruct), ("sky", SkyStruct), ("env", EnvironmentStruct)]


class CameraStruct(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("target", C.c_double * 3), ("fov", C.c_double),
                ("aperture", C.c_double), ("focus_distance", C.c_double)]


TRACE_PER_SAMPLE = 0x1
TRACE_WAVEFRONT = 0x2
TRACE_INLINE_SHADOWS = 0x4


class RenderSettingsStruct(C.Structure):
    _fields_ = [("max_bounces", C.c_uint32), ("firefly_clamp", C.c_float), ("flags", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("occupancy", C.c_uint32), ("tile_w", C.c_uint32),
                ("tile_h", C.c_uint32), ("shade_batch", C.c_uint32), ("trace_turns", C.c_uint32),
                ("job_paths", C.c_uint32)]


class AutotuneTrialStruct(C.Structure):
    _fields_ = [("occupancy", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32),
                ("shade_batch", C.c_uint32), ("trace_turns", C.c_uint32), ("ms", C.c_float),
                ("job_paths", C.c_uint32)]


class BatchItemStruct(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("film", C.c_void_p), ("camera", CameraStruct), ("samples", C.c_uint32)]


class DenoiseParamsStruct(C.Structure):
    _fields_ = [("sigma_l", C.c_float), ("sigma_n", C.c_float), ("sigma_p", C.c_float), ("iterations", C.c_uint32)]


class DisplayParamsStruct(C.Structure):
    _fields_ = [("denoise_weight", C.c_float), ("exposure", C.c_float), ("tonemap", C.c_uint32)]


class GltfSummaryStruct(C.Structure):
    _fields_ = [("mesh_count", C.c_uint32), ("material_count", C.c_uint32), ("image_count", C.c_uint32),
                ("warning_count", C.c_uint32), ("triangle_count", C.c_uint64), ("bounds", C.c_double * 6)]


# name -> (restype, argtypes)
_P = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_D = C.c_double
SIGNATURES: dict[str, tuple] = {
    "lucent_version": (C.c_char_p, []),
    "lucent_last_error": (C.c_char_p, [_P]),
    "lucent_create": (_I32, [C.c_int, C.POINTER(_P)]),
    "lucent_destroy": (None, [_P]),
    "lucent_synchronize": (_I32, [_P]),
    "lucent_device_info": (_I32, [_P, C.POINTER(_I32), C.POINTER(C.c_int64)]),
    "lucent_build_bvh": (_I32, [_P, _U32, _U32, _U32, _P, C.POINTER(_U32), _P, _P]),
    "lucent_build_bvh_gpu": (_I32, [_P, _P, _U32, _U32, _P, C.POINTER(_U32), _P, _P]),
    "lucent_build_bvh_gpu_clustered": (_I32, [_P, _P, _U32, _U32, _U32, _P, C.POINTER(_U32), _P, _P]),
    "lucent_build_bvh_parallel": (_I32, [_P, _U32, _U32, _U32, _U32, _P, C.POINTER(_U32), _P, _P]),
    "lucent_scene_upload": (_I32, [_P, C.POINTER(PackedSceneStruct), C.POINTER(_P)]),
    "lucent_scene_from_bvh": (_I32, [_P, _P, _U32, _P, _U32, _P, C.POINTER(_P)]),
    "lucent_scene_destroy": (None, [_P]),
    "lucent_scene_stats_get": (_I32, [_P, C.POINTER(SceneStatsStruct)]),
    "lucent_intersect": (_I32, [_P, _P, _P, _P, _U32, _D, _D, _U32, _P, _P, _P, _P]),
    "lucent_intersect_f64": (_I32, [_P, _P, _P, _P, _U32, _D, _D, _U32, _P, _P, _P, _P]),
    "lucent_intersect_device": (_I32, [_P, _P, _P, _P, _U32, _D, _D, _U32, _P, _P, _P, _P, _P]),
    "lucent_intersect_device_noalloc": (_I32, [_P, _P, _P, _P, _U32, _D, _D, _U32, _P, _P, _P, _P, _P]),
    "lucent_scene_set_lighting": (_I32, [_P, _P, C.POINTER(LightingStruct)]),
    "lucent_env_sampler": (_I32, [_P, _U32, _U32, _U32, _P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_D)]),
    "lucent_alias_table": (_I32, [_P, _U32, _P, _P, _P, C.POINTER(_D)]),
    "lucent_env_sampler_gpu": (_I32, [_P, _P, _U32, _U32, _U32, _P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_D),
                                      C.POINTER(C.c_float), _P]),
    "lucent_env_sampler_gpu_parallel": (_I32, [_P, _P, _U32, _U32, _U32, _P, C.POINTER(_U32), C.POINTER(_U32),
                                               C.POINTER(_D), C.POINTER(C.c_float), _P]),
    "lucent_scene_rebuild_env_sampler": (_I32, [_P, _P, _U32, C.POINTER(_D)]),
    "lucent_sheen_lut": (_I32, [_U32, _U32, _P]),
    "lucent_film_create": (_I32, [_P, _U32, _U32, C.POINTER(_P)]),
    "lucent_film_destroy": (None, [_P]),
    "lucent_film_reset": (_I32, [_P, _P]),
    "lucent_film_set_seed": (_I32, [_P, _U32]),
    "lucent_film_info": (_I32, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "lucent_film_read": (_I32, [_P, _P, _P, _P, _P, _P]),
    "lucent_film_device_ptrs": (_I32, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "lucent_trace_samples": (_I32, [_P, _P, _P, C.POINTER(CameraStruct), C.POINTER(RenderSettingsStruct), _U32, _P]),
    "lucent_trace_batch": (_I32, [_P, _P, _U32, C.POINTER(RenderSettingsStruct), _P]),
    "lucent_autotune_tiles": (_I32, [_P, _P, _U32, C.POINTER(RenderSettingsStruct), _U32, C.POINTER(_U32), _P]),
    "lucent_autotune_launch": (_I32, [_P, _P, _U32, C.POINTER(RenderSettingsStruct), _U32, _P, _U32,
                                      C.POINTER(_U32)]),
    "lucent_denoise": (_I32, [_P, _P, C.POINTER(CameraStruct), C.POINTER(DenoiseParamsStruct), _P]),
    "lucent_film_read_denoised": (_I32, [_P, _P, _P, _P]),
    "lucent_trace_reprojected": (_I32, [_P, _P, _P, C.POINTER(CameraStruct), _P, C.POINTER(CameraStruct), C.c_float,
                                        C.POINTER(RenderSettingsStruct), _P]),
    "lucent_display": (_I32, [_P, _P, C.POINTER(DisplayParamsStruct), _U32, _U32, _P, _P]),
    "lucent_screenshot": (_I32, [_P, _P, C.POINTER(DisplayParamsStruct), _U32, _U32, _P]),
    "lucent_frame_graph_create": (_I32, [_P, _P, _P, C.POINTER(CameraStruct), C.POINTER(RenderSettingsStruct), _U32,
                                         C.POINTER(DenoiseParamsStruct), C.POINTER(DisplayParamsStruct), _U32, _U32,
                                         _P, C.POINTER(_P)]),
    "lucent_frame_graph_launch": (_I32, [_P, _P]),
    "lucent_batch_graph_create": (_I32, [_P, _P, _U32, C.POINTER(RenderSettingsStruct), C.POINTER(_P)]),
    "lucent_batch_graph_launch": (_I32, [_P, _P]),
    "lucent_batch_graph_destroy": (None, [_P]),
    "lucent_decode_hdr": (_I32, [_P, C.c_size_t, C.POINTER(_U32), C.POINTER(_U32), _P, C.c_size_t]),
    "lucent_gltf_parse": (_I32, [_P, C.c_size_t, C.POINTER(_P)]),
    "lucent_gltf_resource_count": (_U32, [_P]),
    "lucent_gltf_resource_uri": (C.c_char_p, [_P, _U32]),
    "lucent_gltf_set_resource": (_I32, [_P, _U32, _P, C.c_size_t]),
    "lucent_gltf_build": (_I32, [_P]),
    "lucent_gltf_info": (_I32, [_P, C.POINTER(GltfSummaryStruct)]),
    "lucent_gltf_mesh": (_I32, [_P, _U32, C.POINTER(_U32), C.POINTER(_I32), C.POINTER(C.POINTER(C.c_float)),
                                C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.POINTER(C.c_float))]),
    "lucent_gltf_material": (_I32, [_P, _U32, C.POINTER(_D)]),
    "lucent_gltf_image": (_I32, [_P, _U32, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]),
    "lucent_gltf_warning": (C.c_char_p, [_P, _U32]),
    "lucent_gltf_destroy": (None, [_P]),
    "lucent_set_tree_build": (_I32, [_P, _U32, _U32]),
    "lucent_set_record_budget": (_I32, [_P, C.c_uint64]),
    "lucent_frame_graph_destroy": (None, [_P]),
    "lucent_backward": (_I32, [_P, _P, C.POINTER(CameraStruct), C.POINTER(RenderSettingsStruct), _U32, _U32, _U32,
                                _U32, _P, _P, _P, _P, _P]),
    "lucent_forward": (_I32, [_P, _P, _P, C.POINTER(CameraStruct), C.POINTER(RenderSettingsStruct), _U32, _U32, _P,
                               _P]),
    "lucent_detached_forward": (_I32, [_P, _P, C.POINTER(CameraStruct), C.POINTER(RenderSettingsStruct), _U32, _U32,
                                       _U32, _U32, _P, _U32, _P, _P]),
    "lucent_scene_set_materials": (_I32, [_P, _P, _P, _U32]),
    "lucent_scene_set_textures": (_I32, [_P, _P, _U32, _P, _U32, _P, _U32]),
    "lucent_scene_update_env_texels": (_I32, [_P, _P, _P, _U32, _U32]),
    "lucent_scene_env_size": (_I32, [_P, C.POINTER(_U32), C.POINTER(_U32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        # LUCENT_HIP_LIBRARY may point at another build of the same library.
        path = Path(os.environ.get("LUCENT_HIP_LIBRARY", str(LIB_PATH)))
        if not path.exists():
            raise LucentError(-3, f"{path} is missing: run `python -m lucent_amd.build` first")
        l = C.CDLL(str(path), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int) -> None:
    if status != LUCENT_OK:
        msg = lib().lucent_last_error(None)
        raise LucentError(status, msg.decode() if msg else "")
