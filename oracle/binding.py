"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker of the gfx950 library (never as the thing that is
measured or shipped). It takes the same umbreon_amd.Scene / RenderOptions and
marshals them into the same C structs the product consumes.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

from umbreon_amd import capi
from umbreon_amd.render import FrameResult
from umbreon_amd.scene import RenderOptions, Scene, params

ORACLE_LIB = Path(__file__).resolve().parent / "liboracle.so"


class OrcOptions(C.Structure):
    _fields_ = [("precision", C.c_int32), ("threads", C.c_int32), ("row_begin", C.c_int32),
                ("row_end", C.c_int32), ("row_step", C.c_int32), ("brute_force", C.c_int32)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_LIB.exists():
        raise RuntimeError(f"{ORACLE_LIB} is missing; run umbreon_amd/build.py oracle")
    lib = C.CDLL(str(ORACLE_LIB))
    lib.orc_render.argtypes = [C.POINTER(capi.SceneDesc), C.POINTER(capi.RenderParams),
                               C.POINTER(capi.Frame), C.POINTER(OrcOptions)]
    lib.orc_render.restype = C.c_int
    lib.orc_intersect.argtypes = [C.POINTER(capi.SceneDesc), C.c_int64, C.c_void_p, C.c_void_p,
                                  C.POINTER(OrcOptions)]
    lib.orc_intersect.restype = C.c_int
    lib.orc_occluded.argtypes = [C.POINTER(capi.SceneDesc), C.c_int64, C.c_void_p, C.c_void_p,
                                 C.POINTER(OrcOptions)]
    lib.orc_occluded.restype = C.c_int
    lib.orc_occluded_segments.argtypes = [C.POINTER(capi.SceneDesc), C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                          C.c_float, C.c_float, C.c_float, C.c_void_p]
    lib.orc_occluded_segments.restype = C.c_int
    lib.orc_chain_curves.argtypes = [C.POINTER(capi.SceneDesc), C.c_int64, C.c_void_p, C.c_void_p]
    lib.orc_chain_curves.restype = C.c_int64
    lib.orc_object_space_edges.argtypes = [C.POINTER(capi.SceneDesc), C.POINTER(capi.Camera),
                                           C.POINTER(capi.EdgeOptions), C.POINTER(capi.Edges)]
    lib.orc_object_space_edges.restype = C.c_int
    lib.orc_edges_free.argtypes = [C.POINTER(capi.Edges)]
    lib.orc_edges_free.restype = None
    lib.orc_tea2.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_tea2.restype = None
    lib.orc_set_scene_cache.argtypes = [C.c_int32]
    lib.orc_set_scene_cache.restype = None
    lib.orc_fog_factor.argtypes = [C.POINTER(capi.Fog), C.c_float]
    lib.orc_fog_factor.restype = C.c_float
    lib.orc_box_downsample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.orc_box_downsample.restype = None
    lib.orc_srgb_encode.argtypes = [C.c_float]
    lib.orc_srgb_encode.restype = C.c_float
    lib.orc_srgb_decode.argtypes = [C.c_float]
    lib.orc_srgb_decode.restype = C.c_float
    lib.orc_aa_mask.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_float, C.c_float, C.c_int32, C.c_void_p]
    lib.orc_aa_mask.restype = None
    lib.orc_pt1_gather_point.argtypes = [C.POINTER(capi.SceneDesc), C.POINTER(capi.RenderParams), C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_float, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
    lib.orc_pt1_gather_point.restype = C.c_int
    lib.orc_pt1_denoise.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float,
                                    C.c_float, C.c_float]
    lib.orc_pt1_denoise.restype = None
    lib.orc_pt2_sobol_2d.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_pt2_sobol_2d.restype = None
    lib.orc_pt2_pixel_sampler.argtypes = [C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_pt2_pixel_sampler.restype = None
    lib.orc_pathtrace.argtypes = [C.POINTER(capi.SceneDesc), C.POINTER(capi.PtParams), C.c_void_p, C.c_void_p,
                                  C.POINTER(C.c_uint64), C.POINTER(OrcOptions)]
    lib.orc_pathtrace.restype = C.c_int
    lib.orc_pt_bsdf_eval.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_void_p, C.c_void_p]
    lib.orc_pt_bsdf_eval.restype = None
    lib.orc_pt_bsdf_sample.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_pt_bsdf_sample.restype = None
    lib.orc_pt_bsdf_pdf.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
    lib.orc_pt_bsdf_pdf.restype = None
    lib.orc_pt_rnd.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    lib.orc_pt_rnd.restype = C.c_float
    _lib = lib
    return lib


def _opts(precision=32, threads=0, rows=None, brute=False) -> OrcOptions:
    o = OrcOptions()
    o.precision = precision
    o.threads = threads
    if rows is None:
        o.row_begin, o.row_end, o.row_step = 0, 0, 1
    else:
        o.row_begin, o.row_end, o.row_step = rows
    o.brute_force = 1 if brute else 0
    return o


def set_scene_cache(enable: bool) -> None:
    """Reuse the last built scene (geometry + BVH) for renders of an
    unchanged scene (timing only; results are the same)."""
    load().orc_set_scene_cache(1 if enable else 0)


def render(scene: Scene, opt: RenderOptions, hires_only: bool = False, precision: int = 32,
           threads: int = 0, rows=None, brute: bool = False, want_gi_features: bool = False) -> FrameResult:
    lib = load()
    desc, keep1 = scene.to_desc()
    p, keep2 = params(scene, opt, capi.FLAG_HIRES_ONLY if hires_only else 0)
    ss = max(1, int(opt.supersample))
    hw, hh = opt.width * ss, opt.height * ss
    ow, oh = (hw, hh) if hires_only else (opt.width, opt.height)
    color = np.zeros((oh, ow, 4), np.float32)
    depth = np.zeros((hh, hw), np.float32)
    view_z = np.zeros((hh, hw), np.float32)
    hit_id = np.zeros((hh, hw), np.uint32)
    ao = np.zeros((hh, hw), np.float32)
    fr = capi.Frame()
    fr.color, fr.depth, fr.view_z = color.ctypes.data, depth.ctypes.data, view_z.ctypes.data
    fr.hit_id, fr.ao_open = hit_id.ctypes.data, ao.ctypes.data
    adaptive = opt.aa_mode == 1 and (ss > 1 or opt.aa_depth > 1)
    aa_mask = np.zeros((opt.height, opt.width), np.float32) if adaptive else None
    fr.aa_mask = aa_mask.ctypes.data if aa_mask is not None else None
    indirect = np.zeros((hh, hw, 3), np.float32) if opt.gi else None
    gi_occ = np.zeros((hh, hw), np.float32) if opt.gi else None
    fr.gi_indirect = indirect.ctypes.data if indirect is not None else None
    fr.gi_occlusion = gi_occ.ctypes.data if gi_occ is not None else None
    feats = np.zeros((hh, hw, 10), np.float32) if want_gi_features else None
    fr.gi_features = feats.ctypes.data if feats is not None else None
    o = _opts(precision, threads, rows, brute)
    capi.check(lib.orc_render(C.byref(desc), C.byref(p), C.byref(fr), C.byref(o)))
    del keep1, keep2
    return FrameResult(width=ow, height=oh, color=color, depth=depth, view_z=view_z, hit_id=hit_id,
                       ao_open=ao, stats=fr.stats.as_dict(), render_seconds=fr.render_seconds,
                       post_seconds=fr.post_seconds, aa_mask=aa_mask, aa_flagged=int(fr.aa_flagged),
                       indirect=indirect, gi_occlusion=gi_occ, gi_features=feats)


def pt1_gather_point(scene: Scene, opt: RenderOptions, P, N, Ng=None, spp: int = 1024, seed: int = 12345,
                     eps_t: float = 10.0):
    """pt1GatherPoint at one shading point (FP32): returns (E rgb, occlusion, gi rays)."""
    lib = load()
    desc, keep1 = scene.to_desc()
    p, keep2 = params(scene, opt)
    Pa = np.ascontiguousarray(P, np.float32)
    Na = np.ascontiguousarray(N, np.float32)
    Ga = np.ascontiguousarray(N if Ng is None else Ng, np.float32)
    E = np.zeros(3, np.float32)
    occ = C.c_float()
    rays = C.c_uint64()
    capi.check(lib.orc_pt1_gather_point(C.byref(desc), C.byref(p), Pa.ctypes.data, Na.ctypes.data, Ga.ctypes.data,
                                        int(spp), int(seed) & 0xFFFFFFFF, float(eps_t), E.ctypes.data,
                                        C.byref(occ), C.byref(rays)))
    del keep1, keep2
    return E, float(occ.value), int(rays.value)


def aa_mask(rgba, view_z, normal, geom, group, pixel_world_size: float, threshold: float = 0.1,
            perspective: bool = False) -> np.ndarray:
    """buildAaMask (adaptive_aa.hpp:128-194) over hand-built centre results of
    shape (Hf, Wf): rgba (..., 4), view_z, normal (..., 3), geom (uint32,
    0xFFFFFFFF = background), group (int32, -1 = background)."""
    lib = load()
    rgba = np.ascontiguousarray(rgba, np.float32)
    hf, wf = rgba.shape[:2]
    vz = np.ascontiguousarray(view_z, np.float32)
    nr = np.ascontiguousarray(normal, np.float32)
    ge = np.ascontiguousarray(geom, np.uint32)
    gr = np.ascontiguousarray(group, np.int32)
    assert vz.shape == (hf, wf) and nr.shape == (hf, wf, 3) and ge.shape == (hf, wf) and gr.shape == (hf, wf)
    out = np.zeros((hf, wf), np.uint8)
    lib.orc_aa_mask(wf, hf, rgba.ctypes.data, vz.ctypes.data, nr.ctypes.data, ge.ctypes.data, gr.ctypes.data,
                    float(pixel_world_size), float(threshold), int(bool(perspective)), out.ctypes.data)
    return out


def intersect(scene: Scene, rays: np.ndarray, precision: int = 32, brute: bool = False) -> np.ndarray:
    lib = load()
    desc, keep = scene.to_desc()
    rays = np.ascontiguousarray(rays, dtype=capi.ray_dtype())
    hits = np.zeros(len(rays), dtype=capi.hit_dtype())
    o = _opts(precision, 1, None, brute)
    capi.check(lib.orc_intersect(C.byref(desc), len(rays), rays.ctypes.data, hits.ctypes.data, C.byref(o)))
    del keep
    return hits


def occluded(scene: Scene, rays: np.ndarray, precision: int = 32, brute: bool = False) -> np.ndarray:
    lib = load()
    desc, keep = scene.to_desc()
    rays = np.ascontiguousarray(rays, dtype=capi.ray_dtype())
    occ = np.zeros(len(rays), dtype=np.uint8)
    o = _opts(precision, 1, None, brute)
    capi.check(lib.orc_occluded(C.byref(desc), len(rays), rays.ctypes.data, occ.ctypes.data, C.byref(o)))
    del keep
    return occ


def occluded_segments(scene: Scene, segments: np.ndarray, exclude_faces=None, eps: float = 1.0e-4,
                      graze_cos_eps: float = 0.0, coplanar_eps: float = 0.0) -> np.ndarray:
    lib = load()
    desc, keep = scene.to_desc()
    segments = np.ascontiguousarray(segments, dtype=capi.segment_dtype())
    ex = np.ascontiguousarray(np.zeros(0, np.int32) if exclude_faces is None else exclude_faces, dtype=np.int32)
    occ = np.zeros(len(segments), dtype=np.uint8)
    capi.check(lib.orc_occluded_segments(C.byref(desc), len(segments), segments.ctypes.data,
                                         ex.ctypes.data if len(ex) else None, len(ex), eps, graze_cos_eps,
                                         coplanar_eps, occ.ctypes.data))
    del keep
    return occ


def chain_curves(scene: Scene):
    lib = load()
    desc, keep = scene.to_desc()
    cap = max(1, 2 * len(scene.cyl_p0))
    src = np.zeros(cap, np.int32)
    flags = np.zeros(cap, np.uint8)
    n = lib.orc_chain_curves(C.byref(desc), cap, src.ctypes.data, flags.ctypes.data)
    del keep
    assert n >= 0
    return src[:n], flags[:n]


def tea2(v0: int, v1: int):
    lib = load()
    a, b = C.c_uint32(v0), C.c_uint32(v1)
    lib.orc_tea2(C.byref(a), C.byref(b))
    return a.value, b.value


def fog_factor(enabled: bool, start: float, end: float, z: float) -> float:
    lib = load()
    f = capi.Fog()
    f.enabled, f.start, f.end = int(enabled), start, end
    return float(lib.orc_fog_factor(C.byref(f), z))


def box_downsample(src: np.ndarray, ss: int) -> np.ndarray:
    lib = load()
    src = np.ascontiguousarray(src, np.float32)
    h, w, ch = src.shape
    out = np.zeros((h // ss, w // ss, ch), np.float32)
    lib.orc_box_downsample(src.ctypes.data, w, h, ch, ss, out.ctypes.data)
    return out


def srgb_encode(v: float) -> float:
    return float(load().orc_srgb_encode(v))


def srgb_decode(s: float) -> float:
    return float(load().orc_srgb_decode(s))


def pt1_denoise(E, normal, position, iters: int = 5, sigma_z: float = 1.0, sigma_n: float = 128.0,
                sigma_l: float = 4.0) -> np.ndarray:
    """denoisePt1E's a-trous path on (H, W, 3) buffers; returns the filtered E."""
    lib = load()
    E = np.array(E, np.float32, order="C", copy=True)
    h, w = E.shape[:2]
    nr = np.ascontiguousarray(normal, np.float32)
    ps = np.ascontiguousarray(position, np.float32)
    assert nr.shape == E.shape and ps.shape == E.shape
    lib.orc_pt1_denoise(w, h, E.ctypes.data, nr.ctypes.data, ps.ctypes.data, int(iters), float(sigma_z),
                        float(sigma_n), float(sigma_l))
    return E


def pt2_sobol_2d(index: int, seed: int):
    """pt2SobolBurley2D: the 2D Owen-scrambled Sobol point `index` under `seed`."""
    lib = load()
    a, b = C.c_float(), C.c_float()
    lib.orc_pt2_sobol_2d(int(index) & 0xFFFFFFFF, int(seed) & 0xFFFFFFFF, C.byref(a), C.byref(b))
    return a.value, b.value


def pt2_pixel_sampler(pattern: int, px: int, py: int, frame_seed: int, pixel_hash: int, spp: int):
    """pt2MakePixelSampler: (index base, scramble seed) of a gather pixel."""
    lib = load()
    base, seed = C.c_uint32(), C.c_uint32()
    lib.orc_pt2_pixel_sampler(int(pattern), int(px), int(py), int(frame_seed) & 0xFFFFFFFF,
                              int(pixel_hash) & 0xFFFFFFFF, int(spp), C.byref(base), C.byref(seed))
    return base.value, seed.value


def object_space_edges(scene: Scene, opt) -> dict:
    """The oracle's object-space silhouette edges of the spheres and cylinders
    (no mesh feature edges), the union clip against every solid."""
    from umbreon_amd.edges import camera_c, edges_to_numpy
    lib = load()
    desc, keep = scene.to_desc()
    cam, o = camera_c(scene), opt.to_c()
    out = capi.Edges()
    capi.check(lib.orc_object_space_edges(C.byref(desc), C.byref(cam), C.byref(o), C.byref(out)))
    try:
        return edges_to_numpy(out)
    finally:
        lib.orc_edges_free(C.byref(out))
        del keep


def pathtrace(scene: Scene, opt, precision: int = 32, threads: int = 0, brute: bool = False):
    """The path tracer's oracle (umbreon_amd.pathtrace.PtOptions): returns
    (colour H x W x 4, first-hit record H x W x 8, rays traced)."""
    from umbreon_amd.pathtrace import pt_params
    lib = load()
    desc, keep1 = scene.to_desc()
    p, keep2 = pt_params(scene.camera, opt)
    color = np.zeros((opt.height, opt.width, 4), np.float32)
    first = np.zeros((opt.height, opt.width, 8), np.float32)
    rays = C.c_uint64(0)
    o = _opts(precision, threads, None, brute)
    capi.check(lib.orc_pathtrace(C.byref(desc), C.byref(p), color.ctypes.data, first.ctypes.data, C.byref(rays),
                                 C.byref(o)))
    del keep1, keep2
    return color, first, int(rays.value)


def pt_bsdf_eval(mtype: int, roughness: float, ior: float, base, wo, wi) -> np.ndarray:
    """f(wo, wi) |cos wi| (n x 3) of the path tracer's BSDF in the local frame
    (n = +z), FP64, for direction arrays wo, wi (n x 3)."""
    lib = load()
    c = np.ascontiguousarray(np.asarray(base, np.float32).reshape(3))
    a, b = np.broadcast_arrays(np.asarray(wo, np.float64).reshape(-1, 3), np.asarray(wi, np.float64).reshape(-1, 3))
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    out = np.zeros_like(b)
    lib.orc_pt_bsdf_eval(int(mtype), float(roughness), float(ior), c.ctypes.data, len(b), a.ctypes.data,
                         b.ctypes.data, out.ctypes.data)
    return out


def pt_bsdf_sample(mtype: int, roughness: float, ior: float, base, wo, u):
    """BSDF samples for wo (n x 3, or one vector) and uniforms u (n x 3):
    (wi n x 3, weight n x 3 = f |cos| / pdf, ok n)."""
    lib = load()
    c = np.ascontiguousarray(np.asarray(base, np.float32).reshape(3))
    uu = np.ascontiguousarray(np.asarray(u, np.float64).reshape(-1, 3))
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float64).reshape(-1, 3), uu.shape))
    wi = np.zeros_like(uu)
    w = np.zeros_like(uu)
    ok = np.zeros(len(uu), np.uint8)
    lib.orc_pt_bsdf_sample(int(mtype), float(roughness), float(ior), c.ctypes.data, len(uu), a.ctypes.data,
                           uu.ctypes.data, wi.ctypes.data, w.ctypes.data, ok.ctypes.data)
    return wi, w, ok.astype(bool)


def pt_bsdf_pdf(mtype: int, roughness: float, ior: float, base, wo, wi) -> np.ndarray:
    """Density (solid angle, n) with which pt_bsdf_sample draws wi for wo,
    FP64, for direction arrays wo, wi (n x 3)."""
    lib = load()
    c = np.ascontiguousarray(np.asarray(base, np.float32).reshape(3))
    a, b = np.broadcast_arrays(np.asarray(wo, np.float64).reshape(-1, 3), np.asarray(wi, np.float64).reshape(-1, 3))
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    out = np.zeros(len(b))
    lib.orc_pt_bsdf_pdf(int(mtype), float(roughness), float(ior), c.ctypes.data, len(b), a.ctypes.data,
                        b.ctypes.data, out.ctypes.data)
    return out


def pt_rnd(seed: int, pixel: int, sample: int, dim: int) -> float:
    return float(load().orc_pt_rnd(seed, pixel, sample, dim))
