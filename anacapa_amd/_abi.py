"""ctypes mirror of include/anacapa_hip.h (structs and entry points).

Only layout and loading live here; no rendering logic.  The product library is
``anacapa_amd/lib/libanacapa_hip.so``; loading it fails loudly when it has not been
built (there is no CPU fallback for the product path).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libanacapa_hip.so")

AHIP_OK = 0
AHIP_E_INVALID_ARG = 1
AHIP_E_NO_DEVICE = 2
AHIP_E_HIP = 3
AHIP_E_NOT_PREPARED = 4
AHIP_E_UNSUPPORTED = 5
AHIP_E_INJECTED = 6
AHIP_E_OUT_OF_MEMORY = 7

AHIP_MAT_LAMBERT = 0
AHIP_MAT_EMISSIVE = 1
AHIP_MAT_STANDARD = 2
AHIP_MAT_HAIR = 3
AHIP_LIGHT_AREA = 0
AHIP_LIGHT_DOME = 1
AHIP_LIGHT_DIRECTIONAL = 2
AHIP_FLAG_FIX_DOUBLE_COSINE = 1
AHIP_FLAG_HAIR_MARSCHNER = 2
AHIP_INTEGRATOR_PATH = 0
AHIP_INTEGRATOR_BDPT = 1

F3 = C.c_float * 3
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
F32P = C.POINTER(C.c_float)


class Material(C.Structure):
    _fields_ = [
        ("type", C.c_uint32), ("albedo", F3), ("emission", F3), ("emission_weight", C.c_float),
        ("base", C.c_float), ("metalness", C.c_float), ("roughness", C.c_float),
        ("specular", C.c_float), ("specular_color", F3), ("specular_ior", C.c_float),
        ("coat", C.c_float), ("coat_roughness", C.c_float), ("transmission", C.c_float),
        ("roughness_texture", C.c_int32), ("opacity", C.c_float), ("opacity_texture", C.c_int32),
        ("hair_sigma_a", F3), ("hair_eta", C.c_float), ("hair_beta_m", C.c_float), ("hair_beta_n", C.c_float),
        ("hair_alpha", C.c_float), ("alpha_mask", C.c_uint32),
    ]


class Light(C.Structure):
    _fields_ = [
        ("type", C.c_uint32), ("center", F3), ("u_axis", F3), ("v_axis", F3), ("Le", F3),
        ("intensity", C.c_float), ("rot", C.c_float * 9), ("direction", F3), ("half_angle", C.c_float),
        ("scene_center", F3), ("scene_radius", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_uint32), ("mesh_vertex_offset", U32P), ("mesh_index_offset", U32P),
        ("mesh_normal_count", U32P), ("mesh_uv_count", U32P), ("mesh_material", I32P),
        ("positions", F32P), ("normals", F32P), ("uvs", F32P), ("indices", U32P),
        ("num_vertices", C.c_uint32), ("num_indices", C.c_uint32),
        ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)),
        ("num_lights", C.c_uint32), ("lights", C.POINTER(Light)),
        ("env_light", C.c_int32), ("env_radiance", F3),
        ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_pixels", F32P),
        ("num_textures", C.c_uint32), ("texture_width", U32P), ("texture_height", U32P),
        ("texture_pixels", C.POINTER(F32P)),
        ("num_strands", C.c_uint32), ("strand_cv_offset", U32P), ("curve_cvs", F32P), ("strand_width", F32P),
        ("strand_mesh_index", U32P), ("strand_color", F32P),
        ("tri_albedo", F32P),
        ("strand_width_offset", U32P), ("strand_widths", F32P),
    ]


class Camera(C.Structure):
    _fields_ = [("origin", F3), ("lower_left", F3), ("horizontal", F3), ("vertical", F3),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("min_depth", C.c_uint32), ("seed", C.c_uint64),
                ("flags", C.c_uint32), ("curve_depth", C.c_uint32), ("integrator", C.c_uint32),
                ("firefly_clamp", C.c_float)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("sample_start", C.c_uint32), ("sample_count", C.c_uint32),
                ("max_depth", C.c_uint32), ("pad", C.c_uint32), ("accum", F32P)]


class Stats(C.Structure):
    _fields_ = [("rays_closest", C.c_uint64), ("rays_shadow", C.c_uint64), ("samples", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("prims_tested", C.c_uint64),
                ("trace_ms", C.c_double), ("shade_ms", C.c_double),
                ("trace_launches", C.c_uint64), ("shade_launches", C.c_uint64),
                ("render_ms", C.c_double), ("passes", C.c_uint64), ("graph_passes", C.c_uint64),
                ("graph_captures", C.c_uint64)]


AHIP_CURVE_LINEAR, AHIP_CURVE_BEZIER, AHIP_CURVE_BSPLINE, AHIP_CURVE_CATMULL_ROM = 0, 1, 2, 3


class CurveOptions(C.Structure):
    _fields_ = [("default_width", C.c_float), ("width_scale", C.c_float), ("material_index", C.c_uint32),
                ("pad", C.c_uint32)]


class Strands(C.Structure):
    _fields_ = [("num_strands", C.c_uint32), ("num_cvs", C.c_uint32), ("num_skipped", C.c_uint32),
                ("pad", C.c_uint32), ("strand_cv_offset", U32P), ("curve_cvs", F32P),
                ("strand_width_offset", U32P), ("strand_widths", F32P), ("strand_mesh_index", U32P),
                ("strand_color", F32P)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float),
                ("mesh_id", C.c_uint32), ("prim_id", C.c_uint32), ("tri_index", C.c_uint32)]


HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"),
                      ("mesh_id", "<u4"), ("prim_id", "<u4"), ("tri_index", "<u4")])

# Every function the header declares, with its ctypes signature.
VP = C.c_void_p
SIGNATURES = {
    "ahip_create": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "ahip_destroy": (None, [VP]),
    "ahip_is_valid": (C.c_int, [VP]),
    "ahip_device_name": (C.c_char_p, [VP]),
    "ahip_backend_type": (C.c_char_p, []),
    "ahip_last_error": (C.c_char_p, [VP]),
    "ahip_abi_version": (C.c_int, []),
    "ahip_prepare": (C.c_int, [VP, C.POINTER(SceneDesc)]),
    "ahip_bvh_info": (C.c_int, [VP, U32P, U32P, U32P]),
    "ahip_bvh_dump": (C.c_int, [VP, VP, VP]),
    "ahip_set_camera": (C.c_int, [VP, C.POINTER(Camera)]),
    "ahip_set_params": (C.c_int, [VP, C.POINTER(RenderParams)]),
    "ahip_clear_accum": (C.c_int, [VP]),
    "ahip_render_frame": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "ahip_download_accum": (C.c_int, [VP, VP]),
    "ahip_render_tile": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, VP]),
    "ahip_render_batch": (C.c_int, [VP, C.POINTER(View), C.c_uint32]),
    "ahip_resolve": (C.c_int, [VP, C.c_float, VP, VP, VP, VP, VP]),
    "ahip_write_png": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, VP]),
    "ahip_write_exr": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_char_p),
                                 C.POINTER(VP)]),
    "ahip_render_adaptive": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP, VP]),
    "ahip_set_aovs": (C.c_int, [VP, C.c_int]),
    "ahip_download_aovs": (C.c_int, [VP, VP]),
    "ahip_render_tile_aovs": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, VP, VP, VP]),
    "ahip_trace_rays": (C.c_int, [VP, VP, C.c_uint32, VP]),
    "ahip_occluded_rays": (C.c_int, [VP, VP, C.c_uint32, VP]),
    "ahip_trace_rays_noalloc": (C.c_int, [VP, VP, C.c_uint32, VP]),
    "ahip_halton_eval": (C.c_int, [VP, C.c_uint64, VP, VP, VP, VP, C.c_uint32, VP]),
    "ahip_material_probe": (C.c_int, [VP, C.POINTER(Material), C.c_uint32, C.c_uint32, VP, VP]),
    "ahip_set_eval_params": (C.c_int, [VP, VP, C.c_uint32, VP]),
    "ahip_replay": (C.c_int, [VP, C.c_uint32, C.c_uint32, VP]),
    "ahip_adjoint": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, VP, VP, VP]),
    "ahip_get_stats": (C.c_int, [VP, C.POINTER(Stats)]),
    "ahip_reset_stats": (C.c_int, [VP]),
    "ahip_get_bounce_stats": (C.c_int, [VP, C.c_uint32, VP, VP, C.POINTER(C.c_uint32)]),
    "ahip_count_traffic": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "ahip_set_launch": (C.c_int, [VP, C.c_uint32, C.c_uint32]),
    "ahip_fail_next_launch": (C.c_int, [VP]),
    "ahip_camera_pinhole": (C.c_int, [F32P, F32P, F32P, C.c_float, C.c_uint32, C.c_uint32,
                                      C.POINTER(Camera)]),
    "ahip_usd_load": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_char_p, C.POINTER(VP),
                                C.c_char_p, C.c_size_t]),
    "ahip_usd_load_text": (C.c_int, [C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint32, C.c_char_p,
                                     C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "ahip_usd_free": (None, [VP]),
    "ahip_usd_scene_desc": (C.POINTER(SceneDesc), [VP]),
    "ahip_usd_camera": (C.c_int, [VP, F32P, F32P, F32P, F32P, C.POINTER(Camera)]),
    "ahip_usd_mesh_name": (C.c_char_p, [VP, C.c_uint32]),
    "ahip_curves_to_bezier": (C.c_int, [C.c_uint32, C.c_uint32, VP, VP, VP, C.c_uint32, VP, VP,
                                        C.POINTER(CurveOptions), C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "ahip_curve_set_strands": (C.POINTER(Strands), [VP]),
    "ahip_curve_set_free": (None, [VP]),
}

_lib = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load the product library; raises if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"anacapa_amd: native library {p} is not built; run __graft_entry__.build() "
            "(the HIP path has no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    assert a.flags["C_CONTIGUOUS"], "arrays handed to the C ABI must be C-contiguous"
    return C.c_void_p(a.ctypes.data)
