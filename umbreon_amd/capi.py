"""ctypes mirror of the C-ABI in ``include/umbreon_gpu.h``.

The structs below are field-for-field copies of the C declarations; the
loader fails loudly when the gfx950 library has not been built, so nothing in
this package can silently fall back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = Path(__file__).resolve().parent / "lib"
GPU_LIB = LIB_DIR / "libumbreon_gpu.so"

UBG_OK = 0
UBG_ERR_INVALID_ARGUMENT = 1
UBG_ERR_DEVICE = 2
UBG_ERR_OUT_OF_MEMORY = 3
UBG_ERR_CANCELLED = 4
UBG_ERR_UNSUPPORTED = 5
UBG_ERR_INTERNAL = 6

# UBG_ABI_VERSION of include/umbreon_gpu.h this mirror is written against
ABI_VERSION = 8

BVH2 = 0
BVH8 = 1
BVH8_GPU_BUILD = 2  # the BVH8 layout, built on the GPU

KIND_MESH = 0
KIND_SPHERE = 1
KIND_ROUND_CURVE = 2
KIND_CONE_CURVE = 3

FLAG_HIRES_ONLY = 1
FLAG_COUNT_VISITS = 2
FLAG_KERNEL_TIMES = 4
KERNELS = ("primary", "ao", "resolve", "post")

MISS_ID = 0xFFFFFFFF

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u16p = C.POINTER(C.c_uint16)
u8p = C.POINTER(C.c_uint8)


class Material(C.Structure):
    _fields_ = [
        ("ambient", C.c_float), ("diffuse", C.c_float), ("specular", C.c_float),
        ("roughness", C.c_float), ("brilliance", C.c_float), ("phong", C.c_float),
        ("phong_size", C.c_float), ("metallic", C.c_int32), ("reflection", C.c_float),
        ("emission", C.c_float), ("model", C.c_int32), ("pbr_metallic", C.c_float),
        ("pbr_roughness", C.c_float), ("pbr_specular", C.c_float),
        ("pbr_anisotropy", C.c_float), ("pbr_anisotropy_rotation", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int64), ("positions", f32p), ("normals", f32p), ("colors", f32p),
        ("n_triangles", C.c_int64), ("index", u32p),
        ("mesh_materials", C.POINTER(Material)), ("n_mesh_materials", C.c_int32),
        ("tri_material", u8p), ("tri_group", u16p),
        ("n_spheres", C.c_int64), ("sphere_center_radius", f32p), ("sphere_color", f32p),
        ("sphere_material", i32p), ("sphere_group", u16p), ("sphere_from_edge", u8p),
        ("n_cylinders", C.c_int64), ("cyl_p0", f32p), ("cyl_p1", f32p), ("cyl_radius", f32p),
        ("cyl_color", f32p), ("cyl_opacity1", f32p), ("cyl_material", i32p),
        ("cyl_group", u16p), ("cyl_open", u8p), ("cyl_from_edge", u8p),
        ("prim_materials", C.POINTER(Material)), ("n_prim_materials", C.c_int32),
        ("n_instances", C.c_int64), ("instance_offsets", f32p),
    ]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("orthographic", C.c_int32), ("height", C.c_float)]


class Light(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("casts_highlight", C.c_int32), ("angular_radius", C.c_float)]


class Fog(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("color", C.c_float * 3), ("start", C.c_float),
                ("end", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("supersample", C.c_int32),
        ("camera", Camera), ("lights", C.POINTER(Light)), ("n_lights", C.c_int32),
        ("ambient_color", C.c_float * 3), ("background", C.c_float * 3), ("fog", Fog),
        ("clip_near", C.c_float), ("clip_far", C.c_float), ("assumed_gamma", C.c_float),
        ("ao_samples", C.c_int32), ("ao_distance", C.c_float), ("ao_intensity", C.c_float),
        ("shadows", C.c_int32), ("shadow_samples", C.c_int32), ("light_radius_deg", C.c_float),
        ("gi_integrator", C.c_int32), ("specular_scale", C.c_float),
        ("transparency", C.c_int32), ("transparent_background", C.c_int32),
        ("max_transparent_layers", C.c_int32), ("flags", C.c_uint32),
        ("cancel_flag", C.POINTER(C.c_int32)),
        ("ao_falloff_power", C.c_float), ("ao_multi_scale", C.c_int32), ("ao_bent_normal", C.c_int32),
        ("ao_sky_color", C.c_float * 3), ("ao_ground_color", C.c_float * 3), ("ao_use_camera_up", C.c_int32),
        ("ao_up", C.c_float * 3), ("ao_multibounce", C.c_int32), ("ao_low_discrepancy", C.c_int32),
        ("ao_diffuse_factor", C.c_float), ("ao_res_div", C.c_int32),
        ("aa_mode", C.c_int32), ("aa_threshold", C.c_float), ("aa_depth", C.c_int32),
        ("gi", C.c_int32), ("gi_bounces", C.c_int32), ("gi_max_distance", C.c_float),
        ("gi_intensity", C.c_float), ("gi_env_intensity", C.c_float), ("pt1_spp", C.c_int32),
        ("pt1_seed", C.c_uint32), ("pt1_sky_mode", C.c_int32), ("pt1_sky_radiance", C.c_float * 3),
        ("pt1_ld", C.c_int32), ("pt1_clamp", C.c_float), ("pt1_gather_div", C.c_int32),
        ("pt1_half_res", C.c_int32), ("pt1_denoise", C.c_int32), ("denoise_iters", C.c_int32),
        ("denoise_sigma_z", C.c_float), ("denoise_sigma_n", C.c_float), ("denoise_sigma_l", C.c_float),
        ("pt1_upsample_normal_pow", C.c_float), ("pt1_upsample_depth_scale", C.c_float),
        ("pt1_edge_patch", C.c_int32), ("pt1_edge_patch_spp_mul", C.c_int32), ("pt1_edge_patch_thresh", C.c_float),
        ("pt2_pattern", C.c_int32), ("pt2_emissive", C.c_int32), ("pt2_rounds", C.c_int32),
        ("pt2_adaptive", C.c_int32), ("pt2_emissive_nee", C.c_int32), ("pt2_reflect", C.c_int32),
        ("pt1_denoise_net", C.c_void_p),
    ]


class RayStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shaded_hits", C.c_uint64),
                ("ao_rays_traced", C.c_uint64), ("ao_rays_sampled", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("capped_rays", C.c_uint64),
                ("gi_rays", C.c_uint64), ("gi_nee_rays", C.c_uint64), ("gi_gbuffer_rays", C.c_uint64),
                ("node_visits", C.c_uint64 * 3), ("prim_tests", C.c_uint64 * 3)]

    def as_dict(self) -> dict:
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = [int(x) for x in v] if k in ("node_visits", "prim_tests") else int(v)
        return d


class Frame(C.Structure):
    _fields_ = [
        ("color", C.c_void_p), ("depth", C.c_void_p), ("view_z", C.c_void_p),
        ("hit_id", C.c_void_p), ("ao_open", C.c_void_p), ("memory", C.c_int32),
        ("out_w", C.c_int32), ("out_h", C.c_int32), ("hi_w", C.c_int32), ("hi_h", C.c_int32),
        ("stats", RayStats), ("render_seconds", C.c_double), ("post_seconds", C.c_double),
        ("cancelled", C.c_int32), ("layers", C.c_int32), ("kernel_ms", C.c_double * 4),
        ("kernel_launches", C.c_int32 * 4), ("aa_mask", C.c_void_p), ("aa_flagged", C.c_int64),
        ("gi_indirect", C.c_void_p), ("gi_occlusion", C.c_void_p), ("gi_features", C.c_void_p),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("n_triangles", C.c_int64), ("n_spheres", C.c_int64), ("n_round_segments", C.c_int64),
        ("n_cone_segments", C.c_int64), ("n_nodes", C.c_int64), ("n_leaf_prims", C.c_int64),
        ("bvh_depth", C.c_int32), ("node_bytes", C.c_int32), ("bvh_layout", C.c_int32),
        ("stack_depth", C.c_int32), ("prim_bytes", C.c_int32),
        ("build_seconds", C.c_double), ("sah_cost", C.c_double), ("device_bytes", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class EdgeOptions(C.Structure):
    """ubg_edge_options (ObjectSpaceEdgeOptions without `enable`)."""
    _fields_ = [
        ("width", C.c_float), ("raise_", C.c_float), ("segments", C.c_int32), ("color", C.c_float * 3),
        ("clip", C.c_int32), ("mesh_silhouette", C.c_int32), ("mesh_crease", C.c_int32),
        ("mesh_border", C.c_int32), ("visibility_clip", C.c_int32), ("mesh_hard_edge_deg", C.c_float),
        ("crease_angle_deg", C.c_float), ("mesh_crease_smooth_veto_deg", C.c_float),
        ("mesh_crease_convex_only", C.c_int32), ("mesh_border_coplanar_veto_deg", C.c_float),
        ("mesh_crease_max_degree", C.c_int32),
    ]


class Edges(C.Structure):
    """ubg_edges (library-owned arrays)."""
    _fields_ = [
        ("n", C.c_int64), ("p0", C.POINTER(C.c_float)), ("p1", C.POINTER(C.c_float)),
        ("color", C.POINTER(C.c_float)), ("opacity1", C.POINTER(C.c_float)), ("group", C.POINTER(C.c_uint16)),
        ("radius", C.c_float),
    ]


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tnear", C.c_float), ("dir", C.c_float * 3),
                ("tfar", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("kind", C.c_uint32),
                ("prim_id", C.c_uint32), ("ng", C.c_float * 3)]


# numpy structured dtypes matching Ray / Hit (for zero-copy batches)
def segment_dtype():
    """ubg_segment"""
    import numpy as np
    return np.dtype([("p", "<f4", 3), ("q", "<f4", 3), ("exclude_first", "<i4"), ("exclude_count", "<i4")])


def ray_dtype():
    import numpy as np
    return np.dtype([("org", "<f4", 3), ("tnear", "<f4"), ("dir", "<f4", 3), ("tfar", "<f4")])


def hit_dtype():
    import numpy as np
    return np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("kind", "<u4"),
                     ("prim_id", "<u4"), ("ng", "<f4", 3)])


class SceneOptions(C.Structure):
    """ubg_scene_options: BVH layout and leaf size (images do not depend on the leaf size)."""
    _fields_ = [("bvh_layout", C.c_int32), ("leaf_max", C.c_int32), ("reserved", C.c_int32 * 6)]


class BatchOptions(C.Structure):
    """ubg_batch_options: tuning of a batched render (results do not depend on it)."""
    _fields_ = [("ao_waves_per_simd", C.c_int32), ("no_graph", C.c_int32), ("reserved", C.c_int32 * 6)]


class SphereLight(C.Structure):
    """ubg_sphere_light: a spherical area light of uniform radiance."""
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("radiance", C.c_float * 3),
                ("reserved", C.c_float)]


PT_DIFFUSE, PT_CONDUCTOR, PT_DIELECTRIC = 0, 1, 2


class PtMaterial(C.Structure):
    """ubg_pt_material: the path tracer's surface model of one material index."""
    _fields_ = [("type", C.c_int32), ("roughness", C.c_float), ("ior", C.c_float), ("reserved", C.c_int32)]


class PtParams(C.Structure):
    """ubg_pt_params: one path-traced image."""
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("camera", Camera), ("spp", C.c_int32),
        ("bounces", C.c_int32), ("seed", C.c_uint32), ("lights", C.POINTER(SphereLight)), ("n_lights", C.c_int32),
        ("env_radiance", C.c_float * 3), ("prim_materials", C.POINTER(PtMaterial)), ("n_prim_materials", C.c_int32),
        ("mesh_materials", C.POINTER(PtMaterial)), ("n_mesh_materials", C.c_int32),
        ("restir_candidates", C.c_int32), ("restir_spatial", C.c_int32), ("restir_radius", C.c_float),
        ("nee_all_lights", C.c_int32), ("denoise_net", C.c_void_p), ("denoise_ensemble", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class PtFrame(C.Structure):
    """ubg_pt_frame: outputs and timings of ubg_pathtrace."""
    _fields_ = [("color", C.c_void_p), ("features", C.c_void_p), ("memory", C.c_int32),
                ("trace_seconds", C.c_double), ("denoise_seconds", C.c_double), ("rays", C.c_uint64),
                ("reserved", C.c_int32 * 4)]


# Exported symbols of libumbreon_gpu.so (the C-ABI). Kept in sync with the
# header; tests check both directions.
EXPORTS = [
    "ubg_abi_version", "ubg_status_string", "ubg_context_create", "ubg_context_destroy",
    "ubg_last_error", "ubg_scene_create", "ubg_scene_create_ex", "ubg_scene_create_opts", "ubg_scene_destroy", "ubg_scene_get_info",
    "ubg_render", "ubg_render_noalloc", "ubg_render_batch", "ubg_intersect", "ubg_occluded", "ubg_occluded_segments", "ubg_object_space_edges",
    "ubg_edges_free", "ubg_unet_create", "ubg_unet_destroy", "ubg_unet_param_count", "ubg_unet_get_params",
    "ubg_unet_set_params", "ubg_unet_infer", "ubg_unet_train_step", "ubg_unet_set_loss", "ubg_pathtrace",
]

_lib = None


class UmbreonError(RuntimeError):
    """A non-zero status from the C-ABI (the reference throws std::runtime_error)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status


def load() -> C.CDLL:
    """Load libumbreon_gpu.so, raising if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not GPU_LIB.exists():
        raise RuntimeError(
            f"{GPU_LIB} is missing: build the gfx950 backend first "
            "(python -c 'import __graft_entry__; __graft_entry__.build()')")
    lib = C.CDLL(str(GPU_LIB))
    lib.ubg_abi_version.restype = C.c_int
    lib.ubg_status_string.restype = C.c_char_p
    lib.ubg_status_string.argtypes = [C.c_int]
    lib.ubg_last_error.restype = C.c_char_p
    lib.ubg_context_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.ubg_context_create.restype = C.c_int
    lib.ubg_context_destroy.argtypes = [C.c_void_p]
    lib.ubg_context_destroy.restype = None
    lib.ubg_scene_create.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]
    lib.ubg_scene_create.restype = C.c_int
    lib.ubg_scene_create_ex.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.c_int32, C.POINTER(C.c_void_p)]
    lib.ubg_scene_create_ex.restype = C.c_int
    lib.ubg_scene_create_opts.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(SceneOptions),
                                          C.POINTER(C.c_void_p)]
    lib.ubg_scene_create_opts.restype = C.c_int
    lib.ubgx_host_query.argtypes = [C.POINTER(SceneDesc), C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
    lib.ubgx_host_query.restype = C.c_int
    lib.ubg_unet_set_loss.argtypes = [C.c_void_p, C.c_int32]
    lib.ubg_unet_set_loss.restype = C.c_int
    lib.ubgx_unet_set_min_workgroups.argtypes = [C.c_void_p, C.c_int64]
    lib.ubgx_unet_set_min_workgroups.restype = C.c_int
    lib.ubg_scene_destroy.argtypes = [C.c_void_p]
    lib.ubg_scene_destroy.restype = None
    lib.ubg_scene_get_info.argtypes = [C.c_void_p, C.POINTER(SceneInfo)]
    lib.ubg_scene_get_info.restype = C.c_int
    lib.ubg_render.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(RenderParams), C.POINTER(Frame)]
    lib.ubg_render.restype = C.c_int
    lib.ubg_render_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(RenderParams), C.POINTER(Frame)]
    lib.ubg_render_noalloc.restype = C.c_int
    lib.ubg_render_batch.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(RenderParams),
                                     C.POINTER(Frame), C.POINTER(BatchOptions)]
    lib.ubg_render_batch.restype = C.c_int
    lib.ubg_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32]
    lib.ubg_intersect.restype = C.c_int
    lib.ubg_occluded.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32]
    lib.ubg_occluded.restype = C.c_int
    lib.ubg_occluded_segments.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                          C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_int32]
    lib.ubg_occluded_segments.restype = C.c_int
    lib.ubg_object_space_edges.argtypes = [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(Camera),
                                           C.POINTER(EdgeOptions), C.POINTER(Edges)]
    lib.ubg_object_space_edges.restype = C.c_int
    lib.ubg_edges_free.argtypes = [C.POINTER(Edges)]
    lib.ubg_edges_free.restype = None
    lib.ubg_unet_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_uint64, C.POINTER(C.c_void_p)]
    lib.ubg_unet_create.restype = C.c_int
    lib.ubg_unet_destroy.argtypes = [C.c_void_p]
    lib.ubg_unet_destroy.restype = None
    lib.ubg_unet_param_count.argtypes = [C.c_void_p]
    lib.ubg_unet_param_count.restype = C.c_int64
    lib.ubg_unet_get_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.ubg_unet_get_params.restype = C.c_int
    lib.ubg_unet_set_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.ubg_unet_set_params.restype = C.c_int
    lib.ubg_unet_infer.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
    lib.ubg_unet_infer.restype = C.c_int
    lib.ubg_unet_train_step.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_float, C.POINTER(C.c_float), C.c_void_p, C.c_int32]
    lib.ubg_unet_train_step.restype = C.c_int
    lib.ubg_pathtrace.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(PtParams), C.POINTER(PtFrame)]
    lib.ubg_pathtrace.restype = C.c_int
    if lib.ubg_abi_version() != ABI_VERSION:
        raise RuntimeError("libumbreon_gpu.so ABI version mismatch")
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != UBG_OK:
        lib = load()
        msg = lib.ubg_last_error().decode(errors="replace") or lib.ubg_status_string(status).decode()
        raise UmbreonError(status, msg)
