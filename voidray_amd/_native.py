"""ctypes mirror of include/voidray_hip.h and the loader of libvoidray_hip.so.

The library is loaded from this package directory only (built in-tree by
voidray_amd.build).  There is no fallback: if it is missing, every entry
point raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoidray_hip.so")

VR_OK = 0
VR_MAT_LAMBERTIAN, VR_MAT_METAL, VR_MAT_DIELECTRIC, VR_MAT_EMISSION, VR_MAT_LAMBERTIAN_BSDF, VR_MAT_MICROFACET = range(6)
VR_TEX_RGB32F, VR_TEX_RGBA8 = 0, 1
VR_SAMPLE_NEAREST, VR_SAMPLE_BILINEAR = 0, 1
VR_ENV_NONE, VR_ENV_UNIFORM, VR_ENV_HDRI = 0, 1, 2
VR_BUILD_DEFAULT, VR_BUILD_NO_ENV_CDF, VR_BUILD_LEAF_RECORDS, VR_BUILD_LEAF_REFS = 0, 1, 2, 4
VR_ANALYTIC_SPHERE, VR_ANALYTIC_GROUND_PLANE = 0, 1
VR_SURFACE_ANALYTIC = 0x80000000
VR_SURFACE_INSTANCE = 0x40000000
VR_TRACE_CLOSEST, VR_TRACE_ANY, VR_TRACE_TIMING = 0, 1, 0x100
VR_ENGINE_WIDE, VR_ENGINE_REFERENCE, VR_ENGINE_WATERTIGHT = 0, 1, 2
VR_NOALLOC_MAX_RAYS = 1048576
VR_NOALLOC_SCRATCH_BYTES = 75501568
VR_MODE_FULL, VR_MODE_NORMAL = 0, 1
VR_RENDER_ROWS_BY_HEIGHT = 1
VR_RENDER_ENV_IS = 2
VR_RENDER_RUSSIAN_ROULETTE = 4
VR_RENDER_ALPHA_PER_CALL = 8
VR_RENDER_OVERWRITE = 16
VR_BATCH_GRAPH = 1
VR_BATCH_TIMESTAMPS = 2
VR_BATCH_SHORTEST_FIRST = 4
VR_TONEMAP_NONE, VR_TONEMAP_ACES, VR_TONEMAP_REINHARD, VR_TONEMAP_FILMIC, VR_TONEMAP_UNCHARTED2 = range(5)
(VR_TONEMAP_LEGACY_NONE, VR_TONEMAP_LEGACY_SIMPLE_ACES, VR_TONEMAP_LEGACY_SIMPLE_REINHARD,
 VR_TONEMAP_LEGACY_LUMA_REINHARD, VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD,
 VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE, VR_TONEMAP_LEGACY_FILMIC, VR_TONEMAP_LEGACY_UNCHARTED2) = range(16, 24)

F3 = C.c_float * 3


class vr_mesh(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("positions", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p),
                ("indices", C.c_void_p)]


class vr_material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("color", F3), ("param", C.c_float),
                ("albedo_texture", C.c_int32), ("normal_texture", C.c_int32),
                ("index", C.c_float), ("roughness", C.c_float), ("metallic", C.c_float), ("emittance", C.c_float),
                ("transparent", C.c_int32)]


class vr_texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_int32),
                ("sample_type", C.c_int32), ("data", C.c_void_p)]


class vr_environment(C.Structure):
    _fields_ = [("kind", C.c_int32), ("color", F3), ("width", C.c_uint32), ("height", C.c_uint32),
                ("data", C.c_void_p)]


class vr_camera(C.Structure):
    _fields_ = [("eye", F3), ("direction", F3), ("up", F3), ("fov", C.c_float), ("has_dof", C.c_int32),
                ("aperture", C.c_float), ("focal_point", F3)]


class vr_object(C.Structure):
    _fields_ = [("surface", C.c_uint32), ("material", C.c_uint32)]


class vr_analytic(C.Structure):
    _fields_ = [("kind", C.c_int32), ("center", C.c_float * 3), ("radius", C.c_float), ("height", C.c_float)]


class vr_instance(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("transform", C.c_float * 12)]


class vr_scene_desc(C.Structure):
    _fields_ = [("num_meshes", C.c_uint32), ("meshes", C.POINTER(vr_mesh)),
                ("num_surfaces", C.c_uint32), ("surface_mesh", C.c_void_p),
                ("num_objects", C.c_uint32), ("objects", C.POINTER(vr_object)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(vr_material)),
                ("num_textures", C.c_uint32), ("textures", C.POINTER(vr_texture)),
                ("environment", vr_environment), ("camera", vr_camera), ("build_flags", C.c_uint32),
                ("num_analytic", C.c_uint32), ("analytic", C.POINTER(vr_analytic)),
                ("num_instances", C.c_uint32), ("instances", C.POINTER(vr_instance))]


class vr_scene_stats(C.Structure):
    _fields_ = [("num_triangles", C.c_uint32), ("num_surfaces", C.c_uint32), ("num_qnodes", C.c_uint32),
                ("num_ref_nodes", C.c_uint32), ("qbvh_depth", C.c_uint32), ("max_leaf_size", C.c_uint32),
                ("build_ref_seconds", C.c_double), ("build_sah_seconds", C.c_double),
                ("upload_seconds", C.c_double), ("device_bytes", C.c_uint64),
                ("num_tri_records", C.c_uint32), ("leaf_refs", C.c_uint32),
                ("qbox_checked", C.c_uint64), ("qbox_not_containing", C.c_uint64)]


class vr_trace_desc(C.Structure):
    _fields_ = [("num_rays", C.c_uint32), ("t_min", C.c_float), ("mode", C.c_int32), ("engine", C.c_int32)]


class vr_render_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("total_samples", C.c_uint32), ("sample_offset", C.c_uint32), ("max_bounces", C.c_uint32),
                ("firefly_clamp", C.c_float), ("render_mode", C.c_int32), ("flags", C.c_uint32),
                ("rr_depth", C.c_uint32), ("seed", C.c_uint64), ("request", C.c_uint32), ("engine", C.c_uint32)]


class vr_render_stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("samples", C.c_uint64), ("env_misses", C.c_uint64),
                ("fallback_rays", C.c_uint64), ("capped_loops", C.c_uint64), ("kernel_ms", C.c_double),
                ("textured_hits", C.c_uint64), ("env_is_samples", C.c_uint64)]


class vr_post_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("scale", C.c_float), ("gamma", C.c_float),
                ("exposure", C.c_float), ("tonemap", C.c_int32), ("clip_to_1024", C.c_uint32),
                ("keep_alpha", C.c_uint32)]


class vr_request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("desc", vr_render_desc), ("d_rgba", C.c_void_p)]


class vr_tuning(C.Structure):
    _fields_ = [("chunk_paths", C.c_uint32), ("block", C.c_uint32)]


VR_CACHE_BF16 = 0
VR_CACHE_FP32 = 1
VR_CACHE_LEVELS = 16
VR_CACHE_MLP_PARAMS = 21504


class vr_cache_desc(C.Structure):
    _fields_ = [("log2_table", C.c_uint32), ("base_resolution", C.c_uint32), ("max_resolution", C.c_uint32),
                ("precision", C.c_uint32), ("bbox_lo", C.c_float * 3), ("bbox_hi", C.c_float * 3),
                ("lr_grid", C.c_float), ("lr_mlp", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("seed", C.c_uint64)]


class vr_cache_info(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("table_size", C.c_uint32), ("precision", C.c_uint32), ("pad", C.c_uint32),
                ("resolution", C.c_uint32 * 16), ("origin", C.c_float * 3), ("inv_extent", C.c_float),
                ("grid_params", C.c_uint64), ("steps", C.c_uint64)]


class vr_cache_frame(C.Structure):
    _fields_ = [("train_paths", C.c_uint32), ("train_vertices", C.c_uint32), ("cache_depth", C.c_uint32),
                ("train_steps", C.c_uint32), ("frame", C.c_uint32), ("query_roughness", C.c_float)]


class vr_cache_stats(C.Structure):
    _fields_ = [("loss", C.c_double), ("train_samples", C.c_uint64), ("queries", C.c_uint64),
                ("train_ms", C.c_double), ("render_ms", C.c_double)]


# Every symbol the header declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "vr_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "vr_destroy": (C.c_int, [_P]),
    "vr_last_error": (C.c_char_p, []),
    "vr_stream": (_P, [_P]),
    "vr_mesh_load_obj": (C.c_int, [C.c_char_p, C.POINTER(vr_mesh)]),
    "vr_mesh_release": (C.c_int, [C.POINTER(vr_mesh)]),
    "vr_scene_upload": (C.c_int, [_P, C.POINTER(vr_scene_desc), C.POINTER(_P)]),
    "vr_scene_release": (C.c_int, [_P]),
    "vr_scene_get_stats": (C.c_int, [_P, C.POINTER(vr_scene_stats)]),
    "vr_trace": (C.c_int, [_P, C.POINTER(vr_trace_desc), _P, _P, _P, _P, _P, _P, _P]),
    "vr_trace_device_async": (C.c_int, [_P, C.POINTER(vr_trace_desc), _P, _P, _P, _P, _P, _P, _P, _P]),
    "vr_trace_device_noalloc_async": (C.c_int, [_P, C.POINTER(vr_trace_desc), _P, _P, _P, _P, _P, _P, _P, _P]),
    "vr_trace_host": (C.c_int, [C.POINTER(vr_scene_desc), C.POINTER(vr_trace_desc), _P, _P, _P, _P, _P, _P, _P,
                                C.POINTER(vr_scene_stats), C.POINTER(C.c_uint32)]),
    "vr_trace_last_fallbacks": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "vr_trace_last_timing": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "vr_render": (C.c_int, [_P, C.POINTER(vr_render_desc), _P, C.POINTER(vr_render_stats)]),
    "vr_render_device_async": (C.c_int, [_P, C.POINTER(vr_render_desc), _P, C.POINTER(vr_render_stats), _P]),
    "vr_render_last_stats": (C.c_int, [_P, C.POINTER(vr_render_stats)]),
    "vr_render_batch_async": (C.c_int, [_P, C.POINTER(vr_request), C.c_uint32, C.POINTER(vr_tuning), C.c_uint32,
                                        _P]),
    "vr_batch_relaunch_async": (C.c_int, [_P, _P]),
    "vr_batch_latencies": (C.c_int, [_P, C.POINTER(C.c_double), C.c_uint32]),
    "vr_batch_last_stats": (C.c_int, [_P, C.POINTER(vr_render_stats)]),
    "vr_batch_autotune": (C.c_int, [_P, C.POINTER(vr_request), C.c_uint32, C.c_uint32, C.c_uint32, _P,
                                    C.POINTER(vr_tuning), C.POINTER(C.c_float)]),
    "vr_cache_create": (C.c_int, [_P, C.POINTER(vr_cache_desc), C.POINTER(_P)]),
    "vr_cache_destroy": (C.c_int, [_P]),
    "vr_cache_get_info": (C.c_int, [_P, C.POINTER(vr_cache_info)]),
    "vr_cache_get_params": (C.c_int, [_P, _P, _P]),
    "vr_cache_set_params": (C.c_int, [_P, _P, _P]),
    "vr_cache_infer_async": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, _P]),
    "vr_cache_gradients_async": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vr_cache_train_async": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, _P]),
    "vr_render_cached_async": (C.c_int, [_P, _P, C.POINTER(vr_render_desc), C.POINTER(vr_cache_frame), _P, _P]),
    "vr_cache_last_stats": (C.c_int, [_P, C.POINTER(vr_cache_stats)]),
    "vr_tonemap": (C.c_int, [_P, C.POINTER(vr_post_desc), _P, _P]),
    "vr_tonemap_device_async": (C.c_int, [_P, C.POINTER(vr_post_desc), _P, _P, _P]),
}

_lib = None


class VoidrayError(RuntimeError):
    pass


def lib():
    """Loads libvoidray_hip.so from the package directory (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VoidrayError(f"{LIB_PATH} is missing: build it with `python -m voidray_amd.build`")
        # PyTorch-ROCm bundles its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Whichever loads first serves the whole process;
        # if ours came first, torch.cuda would find no device.  Load torch's
        # first when torch is installed, so device tensors and this library
        # share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != VR_OK:
        msg = lib().vr_last_error()
        raise VoidrayError(f"voidray error {rc}: {msg.decode() if msg else ''}")
    return rc
