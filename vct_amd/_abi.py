"""ctypes mirror of include/vct.h and include/vct_app.h.

Struct layouts must match the C headers field for field; tests/test_abi.py
checks the sizes against the compiled library's own view of them.
"""
from __future__ import annotations

import ctypes as C

VCT_OK = 0
VCT_ERR_INVALID = -1
VCT_ERR_HIP = -2
VCT_ERR_STATE = -3
VCT_ERR_NOMEM = -4
VCT_ERR_NODEVICE = -5

VCT_STORE_FP16 = 0
VCT_STORE_FP32 = 1
VCT_STORE_FP8 = 2

VCT_ACCUM_AVG, VCT_ACCUM_MAX, VCT_ACCUM_SUM = 0, 1, 2
VCT_LIGHT_POINT = 0
VCT_LIGHT_DIRECTIONAL = 1

VCT_BUF_ALBEDO = 0
VCT_BUF_NORMAL = 1
VCT_BUF_RADIANCE = 2
VCT_BUF_SHADOWMAP = 3
VCT_BUF_GBUFFER = 4
VCT_BUF_INDIRECT = 5
VCT_BUF_SPECULAR = 6
VCT_BUF_LITMASK = 7
VCT_BUF_RADIANCE_RAW = 8
VCT_BUF_LABELS = 10
VCT_BUF_FEATURES = 11
VCT_BUF_REFERENCE = 12
VCT_BUF_WEIGHTS = 13
VCT_BUF_GRADIENT = 14
VCT_BUF_COLOR = 15
VCT_BUF_WARPMAP = 16
VCT_BUF_WARP_OCCUPANCY = 17
VCT_MLP_PARAMS = 16835
MLP_IN = 160

F16 = C.c_float * 16


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("levels", C.c_int), ("center", C.c_float * 3), ("vmin", C.c_float * 3),
                ("vmax", C.c_float * 3), ("storage", C.c_int), ("anisotropic", C.c_int), ("device", C.c_int)]


class Mesh(C.Structure):
    _fields_ = [("num_tris", C.c_int), ("num_objects", C.c_int), ("positions", C.c_void_p),
                ("normals", C.c_void_p), ("tangents", C.c_void_p), ("albedo", C.c_void_p),
                ("object", C.c_void_p), ("roughness", C.c_void_p), ("vertex_tangents", C.c_void_p),
                ("texcoords", C.c_void_p), ("texture", C.c_void_p)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("levels", C.c_int), ("texels", C.c_void_p)]


class Light(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("range", C.c_float), ("direction", C.c_float * 3),
                ("intensity", C.c_float), ("color", C.c_float * 3), ("type", C.c_uint32),
                ("enabled", C.c_uint32), ("shadow_caster", C.c_uint32), ("shadow_size", C.c_uint32),
                ("perspective", C.c_uint32), ("ls", F16), ("ls_inverse", F16)]


class VoxelizeParams(C.Structure):
    _fields_ = [("mvp_x", F16), ("mvp_y", F16), ("mvp_z", F16), ("axis_override", C.c_int),
                ("lighting", C.c_int), ("atomic_max", C.c_int), ("msaa", C.c_int), ("eye", C.c_float * 3),
                ("warp_texture", C.c_int)]


class InjectParams(C.Structure):
    _fields_ = [("voxel_set_opacity", C.c_float), ("light_mask", C.c_uint32), ("temporal_filter", C.c_int),
                ("temporal_decay", C.c_float), ("warp_texture", C.c_int)]


class MipParams(C.Structure):
    _fields_ = [("kernel_mode", C.c_int), ("albedo_chain", C.c_int)]


class Camera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("eye", C.c_float * 3), ("view", F16),
                ("projection", F16), ("cull_back", C.c_int)]


class ConeSettings(C.Structure):
    _fields_ = [("steps", C.c_int), ("cone_angle", C.c_float), ("bias", C.c_float),
                ("cone_initial_height", C.c_float), ("lod_offset", C.c_float)]


class TraceParams(C.Structure):
    _fields_ = [("diffuse", ConeSettings), ("specular", ConeSettings), ("enable_reflections", C.c_int),
                ("cone_angle_from_roughness", C.c_int), ("draw_occlusion", C.c_int),
                ("ambient_scale", C.c_float), ("reflect_scale", C.c_float), ("eye", C.c_float * 3),
                ("warp_texture", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("total_voxel_fragments", C.c_uint32), ("unique_voxels", C.c_uint32),
                ("max_fragments_per_voxel", C.c_uint32), ("lit_voxels", C.c_uint32)]


class PassTimes(C.Structure):
    _fields_ = [("shadowmap_ms", C.c_double), ("voxelize_ms", C.c_double), ("radiance_ms", C.c_double),
                ("fill_ms", C.c_double), ("mipmap_ms", C.c_double), ("gbuffer_ms", C.c_double),
                ("render_ms", C.c_double), ("learned_ms", C.c_double), ("compose_ms", C.c_double),
                ("warpmap_ms", C.c_double)]


class WarpmapParams(C.Structure):
    _fields_ = [("high_resolution", C.c_float), ("low_resolution", C.c_float), ("axes", C.c_int * 3),
                ("linear", C.c_int), ("use_weights_texture", C.c_int)]


def default_warpmap_params() -> WarpmapParams:
    """Settings defaults (src/Application.h:80-92)."""
    return WarpmapParams(2.0, 0.5, (C.c_int * 3)(1, 1, 1), 0, 1)


class ComposeParams(C.Structure):
    _fields_ = [("cook_torrance", C.c_int), ("enable_diffuse", C.c_int), ("enable_specular", C.c_int),
                ("enable_indirect", C.c_int), ("enable_reflections", C.c_int), ("postprocess", C.c_int),
                ("ambient_scale", C.c_float), ("reflect_scale", C.c_float), ("shininess", C.c_float),
                ("eye", C.c_float * 3)]


class AppConfig(C.Structure):
    _fields_ = [("scene", C.c_int), ("seed", C.c_uint64), ("dim", C.c_int), ("levels", C.c_int),
                ("storage", C.c_int), ("anisotropic", C.c_int), ("width", C.c_int), ("height", C.c_int),
                ("shadow_size", C.c_int), ("lighting", C.c_int), ("msaa", C.c_int), ("reflections", C.c_int),
                ("fill_holes", C.c_int), ("device", C.c_int), ("temporal_filter", C.c_int),
                ("temporal_decay", C.c_float), ("track_camera", C.c_int), ("warp_texture", C.c_int),
                ("warp_high_resolution", C.c_float), ("tessellation", C.c_int)]


class LearnedParams(C.Structure):
    _fields_ = [("batch", C.c_int), ("steps", C.c_int), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint32), ("reset", C.c_int),
                ("lr_final", C.c_float)]


GBUFFER_TEXEL_BYTES = 64

# (name, restype, argtypes) of every function include/vct.h declares
VCT_FUNCS = [
    ("vct_abi_version", C.c_int, []),
    ("vct_last_error", C.c_char_p, [C.c_void_p]),
    ("vct_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(GridDesc)]),
    ("vct_remake", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("vct_destroy", None, [C.c_void_p]),
    ("vct_get_grid", C.c_int, [C.c_void_p, C.POINTER(GridDesc)]),
    ("vct_set_extent", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vct_set_scene", C.c_int, [C.c_void_p, C.POINTER(Mesh)]),
    ("vct_set_textures", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("vct_set_transforms", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("vct_set_lights", C.c_int, [C.c_void_p, C.POINTER(Light), C.c_int]),
    ("vct_shadowmap", C.c_int, [C.c_void_p, C.c_int]),
    ("vct_shadowmaps", C.c_int, [C.c_void_p]),
    ("vct_voxelize", C.c_int, [C.c_void_p, C.POINTER(VoxelizeParams)]),
    ("vct_voxelize_tessellated", C.c_int, [C.c_void_p, C.POINTER(VoxelizeParams)]),
    ("vct_normalize_inject", C.c_int, [C.c_void_p, C.POINTER(InjectParams)]),
    ("vct_normalize_voxels", C.c_int, [C.c_void_p, C.c_float]),
    ("vct_normalize_inject_noalloc", C.c_int, [C.c_void_p, C.POINTER(InjectParams)]),
    ("vct_fill_holes", C.c_int, [C.c_void_p]),
    ("vct_mip", C.c_int, [C.c_void_p, C.POINTER(MipParams)]),
    ("vct_gbuffer", C.c_int, [C.c_void_p, C.POINTER(Camera)]),
    ("vct_trace", C.c_int, [C.c_void_p, C.POINTER(TraceParams)]),
    ("vct_compose", C.c_int, [C.c_void_p, C.POINTER(ComposeParams)]),
    ("vct_warpmap", C.c_int, [C.c_void_p, C.POINTER(VoxelizeParams), C.POINTER(WarpmapParams)]),
    ("vct_sync", C.c_int, [C.c_void_p]),
    ("vct_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("vct_get_pass_times", C.c_int, [C.c_void_p, C.POINTER(PassTimes)]),
    ("vct_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("vct_buffer_bytes", C.c_size_t, [C.c_void_p, C.c_int, C.c_int]),
    ("vct_download", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
    ("vct_upload", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
    ("vct_learned_set_points", C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ("vct_learned_prepare", C.c_int, [C.c_void_p]),
    ("vct_learned_train", C.c_int, [C.c_void_p, C.POINTER(LearnedParams), C.c_void_p]),
    ("vct_trace_learned", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.c_float, C.POINTER(C.c_int)]),
    ("vct_reference_gather", C.c_int, [C.c_void_p]),
]

APP_FUNCS = [
    ("vct_app_default_config", C.c_int, [C.c_int, C.POINTER(AppConfig)]),
    ("vct_app_create", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(AppConfig)]),
    ("vct_app_create_from_obj", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(AppConfig), C.c_char_p]),
    ("vct_obj_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    ("vct_obj_mesh", C.c_int, [C.c_void_p, C.POINTER(Mesh)]),
    ("vct_obj_bounds", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    ("vct_obj_textures", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    ("vct_obj_free", None, [C.c_void_p]),
    ("vct_dds_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    ("vct_png_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    ("vct_dds_texture", C.c_int, [C.c_void_p, C.POINTER(Texture)]),
    ("vct_dds_free", None, [C.c_void_p]),
    ("vct_app_destroy", None, [C.c_void_p]),
    ("vct_app_last_error", C.c_char_p, [C.c_void_p]),
    ("vct_app_ctx", C.c_void_p, [C.c_void_p]),
    ("vct_app_get_config", C.c_int, [C.c_void_p, C.POINTER(AppConfig)]),
    ("vct_app_scene_sizes", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("vct_app_get_mesh", C.c_int, [C.c_void_p] + [C.c_void_p] * 6),
    ("vct_app_get_vertex_tangents", C.c_int, [C.c_void_p, C.c_void_p]),
    ("vct_app_frame_state", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(Light),
                                      C.POINTER(VoxelizeParams), C.POINTER(InjectParams), C.POINTER(MipParams),
                                      C.POINTER(Camera), C.POINTER(TraceParams)]),
    ("vct_app_render", C.c_int, [C.c_void_p, C.c_int]),
    ("vct_app_compose_params", C.c_int, [C.c_void_p, C.c_int, C.POINTER(ComposeParams)]),
    ("vct_app_warpmap_params", C.c_int, [C.c_void_p, C.POINTER(WarpmapParams)]),
    ("vct_app_build_radiance", C.c_int, [C.c_void_p, C.c_int]),
    ("vct_app_surface_points", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_void_p]),
    ("vct_app_camera", C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.POINTER(Camera)]),
    ("vct_math_ortho", None, [C.c_float] * 6 + [C.c_void_p]),
    ("vct_math_lookat", None, [C.c_void_p] * 4),
    ("vct_math_perspective", None, [C.c_float] * 4 + [C.c_void_p]),
    ("vct_math_inverse4", C.c_int, [C.c_void_p, C.c_void_p]),
    ("vct_math_mul4", None, [C.c_void_p] * 3),
    ("vct_math_voxel_projections", None, [C.c_void_p] * 6),
]


def bind(lib: C.CDLL, funcs) -> None:
    for name, res, args in funcs:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
