"""ctypes binding of libhdsr.so (the C ABI declared in include/hdsr.h).

The library is the only compute path: there is no CPU fallback, and loading fails loudly when
the built library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# HDSR_LIB_PATH selects another build of the library (for comparing builds); default: in-tree
LIB_PATH = os.environ.get("HDSR_LIB_PATH") or os.path.join(_PKG, "lib", "libhdsr.so")

HDSR_OK = 0
HDSR_E_INVALID_ARG = -1
HDSR_E_OOM = -2
HDSR_E_CAPACITY = -3
HDSR_E_HIP = -4
HDSR_E_STATE = -5
HDSR_E_UNSUPPORTED = -6

HDSR_F16_COMPOSITE = 1
HDSR_REF_CAPS = 2
HDSR_GBUFFER = 4
HDSR_DEBUG = 8
HDSR_OUT_RGBA16 = 16
HDSR_PROFILE = 32
HDSR_DIFFERENTIABLE = 64
HDSR_BUCKET_SCATTER = 128
STAGES = ["project", "scan", "emit", "group", "ranges", "classify", "bucket", "sort", "composite"]

HDSR_MEM_HOST = 0
HDSR_MEM_DEVICE = 1
HDSR_SH_DEGREE_AUTO = 0xFFFFFFFF  # take the SH degree from the shData length
HDSR_NOALLOC_MAX_SPLATS = 4194304  # scratch capacity of hdsr_sort_back_to_front_noalloc

(OUT_RGBA, OUT_DEPTH, OUT_GB_NORMAL, OUT_GB_MATERIAL, OUT_RGBA16, OUT_T_FINAL, OUT_LAST_INDEX, OUT_FILTERED_NORMAL,
 OUT_LIT) = range(9)

DBG = {
    "projcache": 0, "depth": 1, "tile_counts": 2, "tile_offsets": 3, "entries": 4, "entries_canon": 5,
    "small_list": 6, "large_list": 7, "large_bases": 8, "chunks": 9, "keys20": 10, "large_splats": 11, "sorted": 12,
}


class PlanIn(C.Structure):
    _fields_ = [("viewport_width", C.c_uint32), ("viewport_height", C.c_uint32), ("tile_size_px", C.c_uint32),
                ("splat_count", C.c_uint32), ("tile_entry_multiplier", C.c_float)]


class Plan(C.Structure):
    _fields_ = [("viewport_width", C.c_uint32), ("viewport_height", C.c_uint32), ("tile_size_px", C.c_uint32),
                ("tile_columns", C.c_uint32), ("tile_rows", C.c_uint32), ("tile_count", C.c_uint32),
                ("morton_tile_count", C.c_uint32), ("splat_count", C.c_uint32),
                ("max_total_tile_refs", C.c_uint32)]


class Splats(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p),
                ("opacities", C.c_void_p), ("materials", C.c_void_p), ("normals", C.c_void_p),
                ("normals_len", C.c_uint64), ("sh_data", C.c_void_p), ("sorted_indices", C.c_void_p),
                ("count", C.c_uint32), ("sh_degree", C.c_uint32), ("memory", C.c_uint32),
                ("sh_data_len", C.c_uint64)]


class Outputs(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("depth", C.c_void_p), ("gbuffer_normal", C.c_void_p),
                ("gbuffer_material", C.c_void_p), ("rgba16", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("lit", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "refs", "small_pairs", "large_pairs", "large_splats", "visible_splats", "small_tiles", "large_tiles",
        "chunks", "segments", "pair_capacity", "overflow_reruns", "cap_refs_tripped", "cap_large_splats_tripped",
        "cap_chunks_tripped", "composite_refs", "composite_max_refs", "large_tile_refs", "cell_pairs",
        "sorted_refs", "resumed_tiles", "resumed_tile_refs", "first_chunks")]


class Grads(C.Structure):
    _fields_ = [("means", C.c_void_p), ("log_scales", C.c_void_p), ("rotations", C.c_void_p),
                ("opacity_logits", C.c_void_p), ("sh_data", C.c_void_p)]


class NormalParams(C.Structure):
    _fields_ = [("recovery_mode", C.c_uint32), ("grazing_start", C.c_float), ("grazing_end", C.c_float),
                ("spatial_sigma", C.c_float), ("depth_sigma", C.c_float), ("normal_sigma", C.c_float),
                ("kernel_radius", C.c_uint32)]


class MaterialCurve(C.Structure):
    _fields_ = [("contrast", C.c_double), ("brightness", C.c_double), ("gamma", C.c_double)]


class LightingParams(C.Structure):
    _fields_ = [("light_direction", C.c_double * 3), ("light_intensity", C.c_double),
                ("ambient_intensity", C.c_double), ("specular_only", C.c_uint32),
                ("emissive_intensity", C.c_double), ("emissive_threshold", C.c_double),
                ("env_intensity", C.c_double), ("env_rotation", C.c_double), ("exposure", C.c_double),
                ("roughness_curve", MaterialCurve), ("metalness_curve", MaterialCurve),
                ("albedo_curve", MaterialCurve), ("lighting_view_proj", C.c_void_p),
                ("lighting_camera_position", C.c_void_p), ("ao", C.c_void_p), ("bloom", C.c_void_p)]


class PlyInfo(C.Structure):
    _fields_ = [("count", C.c_uint32), ("sh_degree", C.c_uint32), ("sh_coefficient_count", C.c_uint32),
                ("fields", C.c_uint32), ("data_offset", C.c_uint64), ("row_stride", C.c_uint32),
                ("property_count", C.c_uint32)]


class SplatAttributesOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("positions", "colors", "opacities", "radii", "scales", "rotations",
                                          "sh_coefficients", "normals", "roughness", "metalness", "emissive",
                                          "detail_normals", "original_ids")] + [
        ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3), ("bounds_center", C.c_double * 3),
        ("bounds_radius", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_splats_per_batch", C.c_uint64),
                ("max_views_per_batch", C.c_uint32), ("slots", C.c_uint32), ("use_graphs", C.c_uint32),
                ("reserved", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("scene_id", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("reserved", C.c_uint32),
                ("uniforms256", C.c_void_p)]


class ResultInfo(C.Structure):
    _fields_ = [("done", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("batch_views", C.c_uint32),
                ("latency_ms", C.c_double), ("rgba", C.c_void_p), ("depth", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("batches", C.c_uint64), ("graph_replays", C.c_uint64),
                ("uploads", C.c_uint64), ("evictions", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("resident_scenes", C.c_uint32), ("queued", C.c_uint32), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_batch_views", C.c_double), ("busy_s", C.c_double)]


# Every symbol include/hdsr.h and include/hdsr_assets.h declare: (name, restype, argtypes).
SIGNATURES = [
    ("hdsr_plan_compute", C.c_int, [C.POINTER(PlanIn), C.POINTER(Plan)]),
    ("hdsr_create", C.c_int, [C.POINTER(Plan), C.c_uint32, C.POINTER(C.c_void_p)]),
    ("hdsr_destroy", None, [C.c_void_p]),
    ("hdsr_bind_splats", C.c_int, [C.c_void_p, C.POINTER(Splats)]),
    ("hdsr_update_sorted_indices", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    ("hdsr_set_frame", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hdsr_render", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hdsr_sort_back_to_front", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hdsr_sort_back_to_front_noalloc", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("hdsr_composite_only", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hdsr_sync", C.c_int, [C.c_void_p]),
    ("hdsr_get_outputs", C.c_int, [C.c_void_p, C.POINTER(Outputs)]),
    ("hdsr_bind_outputs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("hdsr_readback", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]),
    ("hdsr_debug_size", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]),
    ("hdsr_debug_readback", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]),
    ("hdsr_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("hdsr_stage_times", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
    ("hdsr_backward", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Grads), C.c_void_p]),
    ("hdsr_normal_params_default", None, [C.POINTER(NormalParams)]),
    ("hdsr_lighting_params_default", None, [C.POINTER(LightingParams)]),
    ("hdsr_deferred_lighting", C.c_int, [C.c_void_p, C.POINTER(LightingParams), C.c_void_p]),
    ("hdsr_write_lighting_params", None, [C.POINTER(LightingParams), C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_void_p]),
    ("hdsr_material_curve_lut", None, [C.POINTER(MaterialCurve), C.c_void_p]),
    ("hdsr_screen_space_normals", C.c_int, [C.c_void_p, C.POINTER(NormalParams), C.c_void_p]),
    ("hdsr_bilateral_normal_filter", C.c_int, [C.c_void_p, C.POINTER(NormalParams), C.c_void_p]),
    ("hdsr_ply_inspect", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(PlyInfo), C.c_char_p, C.c_uint32]),
    ("hdsr_ply_decode", C.c_int, [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(SplatAttributesOut), C.c_char_p,
                                  C.c_uint32]),
    ("hdsr_last_error", C.c_char_p, [C.c_void_p]),
    ("hdsr_version", C.c_char_p, []),
    ("hdsr_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    ("hdsr_server_destroy", None, [C.c_void_p]),
    ("hdsr_server_add_scene", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Splats)]),
    ("hdsr_server_remove_scene", C.c_int, [C.c_void_p, C.c_uint32]),
    ("hdsr_server_submit", C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    ("hdsr_server_wait", C.c_int, [C.c_void_p, C.c_uint64, C.c_double]),
    ("hdsr_server_result", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(ResultInfo)]),
    ("hdsr_server_read", C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint64]),
    ("hdsr_server_release", C.c_int, [C.c_void_p, C.c_uint64]),
    ("hdsr_server_drain", C.c_int, [C.c_void_p]),
    ("hdsr_server_hold", C.c_int, [C.c_void_p, C.c_uint32]),
    ("hdsr_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    ("hdsr_server_reset_stats", C.c_int, [C.c_void_p]),
    ("hdsr_server_last_error", C.c_char_p, [C.c_void_p]),
]

_lib = None


def load() -> C.CDLL:
    """Loads libhdsr.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP compositor library not built: {LIB_PATH} (run build.py)")
    # One HIP runtime per process.  torch ships its own libamdhip64 and loads it by path; loaded
    # before torch, libhdsr would map /opt/rocm's copy as well, and of two runtimes the one that
    # initialises second finds no device (hipErrorNoDevice).  With torch imported first, libhdsr's
    # libamdhip64.so.7 dependency binds to torch's copy (same soname), so torch's device pointers
    # and streams are the library's too.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HdsrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hdsr error {code}: {msg}")
        self.code = code


def check(rc: int, ctx=None) -> None:
    if rc != HDSR_OK:
        msg = ""
        if ctx is not None:
            m = load().hdsr_last_error(ctx)
            msg = m.decode() if m else ""
        raise HdsrError(rc, msg)
