"""Host-side driver over the C-ABI: device context, built scene, frame.

Mirrors the reference's public flow -- ``umbreon::render(scene, opt) ->
FrameResult`` (src/umbreon/umbreon.hpp:29) -- with the device context and the
built scene exposed so a caller can reuse them across frames (the
``SharedDevice`` reuse of src/umbreon/umbreon.cpp:21-39).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import capi
from .scene import RenderOptions, Scene, params


@dataclass
class FrameResult:
    """umbreon::FrameResult (render/frame_result.hpp:43-124), numpy-backed."""
    width: int
    height: int
    color: np.ndarray                  # (H, W, 4) linear HDR RGBA
    depth: Optional[np.ndarray]        # (hiH, hiW) first-hit distance (0 on miss)
    view_z: Optional[np.ndarray]       # (hiH, hiW) linear view-z
    hit_id: Optional[np.ndarray] = None
    ao_open: Optional[np.ndarray] = None
    stats: dict = field(default_factory=dict)
    render_seconds: float = 0.0
    post_seconds: float = 0.0
    cancelled: bool = False
    layers: int = 0
    kernel_ms: dict = field(default_factory=dict)
    kernel_launches: dict = field(default_factory=dict)
    aa_mask: Optional[np.ndarray] = None  # (H, W) 1 = refined (adaptive AA with aa_debug)
    aa_flagged: int = 0
    indirect: Optional[np.ndarray] = None      # (hiH, hiW, 3) GI indirect added to colour
    gi_occlusion: Optional[np.ndarray] = None  # (hiH, hiW) first-bounce hit fraction
    gi_features: Optional[np.ndarray] = None   # (hiH, hiW, 10) learned-denoiser features of the gather

    @property
    def total_rays(self) -> int:
        """Every ray of the frame: primary (with continuations), AO (traced and
        sampled), shadow, and the GI gather, next-event and G-buffer rays."""
        s = self.stats
        return int(s["primary_rays"] + s["ao_rays_traced"] + s["ao_rays_sampled"] + s["shadow_rays"]
                   + s.get("gi_rays", 0) + s.get("gi_nee_rays", 0) + s.get("gi_gbuffer_rays", 0))


class Device:
    """One gfx950 device context (stream + reusable frame scratch)."""

    def __init__(self, ordinal: int = 0):
        self._lib = capi.load()
        h = C.c_void_p()
        capi.check(self._lib.ubg_context_create(int(ordinal), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self._lib.ubg_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BuiltScene:
    """Scene geometry resident in HBM with its BVH (buildEmbreeScene analogue)."""

    def __init__(self, device: Device, scene: Scene, layout: int = capi.BVH8, leaf_max: int = 0):
        self.device = device
        self._lib = device._lib
        desc, keep = scene.to_desc()
        h = C.c_void_p()
        so = capi.SceneOptions()
        so.bvh_layout, so.leaf_max = int(layout), int(leaf_max)
        capi.check(self._lib.ubg_scene_create_opts(device.handle, C.byref(desc), C.byref(so), C.byref(h)))
        del keep
        self.handle = h
        info = capi.SceneInfo()
        capi.check(self._lib.ubg_scene_get_info(self.handle, C.byref(info)))
        self.info = info.as_dict()

    def close(self) -> None:
        if self.handle:
            self._lib.ubg_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # rtcIntersect1 / rtcOccluded1, batched
    def intersect(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=capi.ray_dtype())
        hits = np.zeros(len(rays), dtype=capi.hit_dtype())
        capi.check(self._lib.ubg_intersect(self.device.handle, self.handle, len(rays),
                                           rays.ctypes.data, hits.ctypes.data, 0))
        return hits

    def occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=capi.ray_dtype())
        occ = np.zeros(len(rays), dtype=np.uint8)
        capi.check(self._lib.ubg_occluded(self.device.handle, self.handle, len(rays),
                                          rays.ctypes.data, occ.ctypes.data, 0))
        return occ

    def occluded_segments(self, segments: np.ndarray, exclude_faces=None, eps: float = 1.0e-4,
                          graze_cos_eps: float = 0.0, coplanar_eps: float = 0.0) -> np.ndarray:
        """EmbreeRenderer::occluded over a batch (ubg_occluded_segments)."""
        segments = np.ascontiguousarray(segments, dtype=capi.segment_dtype())
        ex = np.ascontiguousarray(np.zeros(0, np.int32) if exclude_faces is None else exclude_faces, dtype=np.int32)
        occ = np.zeros(len(segments), dtype=np.uint8)
        capi.check(self._lib.ubg_occluded_segments(self.device.handle, self.handle, len(segments),
                                                   segments.ctypes.data, ex.ctypes.data if len(ex) else None,
                                                   len(ex), eps, graze_cos_eps, coplanar_eps, occ.ctypes.data, 0))
        return occ

    def render(self, scene: Scene, opt: RenderOptions, hires_only: bool = False,
               want_ids: bool = False, want_ao: bool = False, count_visits: bool = False,
               kernel_times: bool = False, outputs: bool = True,
               cancel_flag: Optional[C.c_int32] = None, want_gi_features: bool = False,
               denoise_net=None, color_only: bool = False) -> FrameResult:
        """One frame. outputs=False leaves the frame in HBM (no host copy);
        color_only copies the colour alone (depth, view z and the GI AOVs
        stay on the device).
        want_gi_features returns the learned denoiser's input features of the
        gathered irradiance (hi-res H x W x 10); denoise_net (an
        umbreon_amd.denoise.Unet) runs the learned denoiser in place of the
        a-trous filter (opt.pt1_denoise is then 2)."""
        flags = ((capi.FLAG_HIRES_ONLY if hires_only else 0) | (capi.FLAG_COUNT_VISITS if count_visits else 0)
                 | (capi.FLAG_KERNEL_TIMES if kernel_times else 0))
        p, keep = params(scene, opt, flags, cancel_flag)
        if denoise_net is not None:
            p.pt1_denoise = 2
            p.pt1_denoise_net = denoise_net.handle
        ss = max(1, int(opt.supersample))
        hw, hh = opt.width * ss, opt.height * ss
        ow, oh = (hw, hh) if hires_only else (opt.width, opt.height)
        color = (np.empty if color_only else np.zeros)((oh, ow, 4), np.float32) if outputs else None
        aovs = outputs and not color_only
        depth = np.zeros((hh, hw), np.float32) if aovs else None
        view_z = np.zeros((hh, hw), np.float32) if aovs else None
        hit_id = np.zeros((hh, hw), np.uint32) if want_ids else None
        ao = np.zeros((hh, hw), np.float32) if want_ao else None
        fr = capi.Frame()
        fr.color = color.ctypes.data if color is not None else None
        fr.depth = depth.ctypes.data if depth is not None else None
        fr.view_z = view_z.ctypes.data if view_z is not None else None
        fr.hit_id = hit_id.ctypes.data if hit_id is not None else None
        fr.ao_open = ao.ctypes.data if ao is not None else None
        adaptive = opt.aa_mode == 1 and (ss > 1 or opt.aa_depth > 1)
        aa_mask = np.zeros((opt.height, opt.width), np.float32) if opt.aa_debug and adaptive else None
        fr.aa_mask = aa_mask.ctypes.data if aa_mask is not None else None
        indirect = np.zeros((hh, hw, 3), np.float32) if opt.gi and aovs else None
        gi_occ = np.zeros((hh, hw), np.float32) if opt.gi and aovs else None
        fr.gi_indirect = indirect.ctypes.data if indirect is not None else None
        fr.gi_occlusion = gi_occ.ctypes.data if gi_occ is not None else None
        feats = np.zeros((hh, hw, 10), np.float32) if want_gi_features else None
        fr.gi_features = feats.ctypes.data if feats is not None else None
        fr.memory = 0
        capi.check(self._lib.ubg_render(self.device.handle, self.handle, C.byref(p), C.byref(fr)))
        del keep
        return FrameResult(width=ow, height=oh, color=color, depth=depth, view_z=view_z,
                           hit_id=hit_id, ao_open=ao, stats=fr.stats.as_dict(),
                           render_seconds=fr.render_seconds, post_seconds=fr.post_seconds,
                           cancelled=bool(fr.cancelled), layers=int(fr.layers),
                           kernel_ms={k: float(fr.kernel_ms[i]) for i, k in enumerate(capi.KERNELS)},
                           kernel_launches={k: int(fr.kernel_launches[i]) for i, k in enumerate(capi.KERNELS)},
                           aa_mask=aa_mask, aa_flagged=int(fr.aa_flagged), indirect=indirect,
                           gi_occlusion=gi_occ, gi_features=feats)


def render_batch(device: Device, built: Sequence["BuiltScene"], scenes: Sequence[Scene],
                 opts: Sequence[RenderOptions], want_ids: bool = False, want_ao: bool = False,
                 hires_only: bool = False, ao_waves_per_simd: int = 0, color_only: bool = False,
                 graph: bool = True) -> list:
    """Several requests in one batched render (ubg_render_batch): request i
    is scene scenes[i] (built as built[i]) with options opts[i]; returns one
    FrameResult per request, each equal to BuiltScene.render of it alone
    (color_only: only the colour comes back to the host)."""
    n = len(built)
    assert len(scenes) == n and len(opts) == n and n > 0
    lib = device._lib
    keeps, outs = [], []
    P = (capi.RenderParams * n)()
    F = (capi.Frame * n)()
    H = (C.c_void_p * n)(*[b.handle.value for b in built])
    for i in range(n):
        flags = capi.FLAG_HIRES_ONLY if hires_only else 0
        p, keep = params(scenes[i], opts[i], flags)
        P[i] = p
        keeps.append(keep)
        opt = opts[i]
        ss = max(1, int(opt.supersample))
        hw, hh = opt.width * ss, opt.height * ss
        ow, oh = (hw, hh) if hires_only else (opt.width, opt.height)
        o = {"color": np.empty((oh, ow, 4), np.float32),
             "depth": None if color_only else np.zeros((hh, hw), np.float32),
             "view_z": None if color_only else np.zeros((hh, hw), np.float32),
             "hit_id": np.zeros((hh, hw), np.uint32) if want_ids else None,
             "ao_open": np.zeros((hh, hw), np.float32) if want_ao else None, "w": ow, "h": oh}
        outs.append(o)
        F[i].color = o["color"].ctypes.data
        F[i].depth = None if color_only else o["depth"].ctypes.data
        F[i].view_z = None if color_only else o["view_z"].ctypes.data
        F[i].hit_id = o["hit_id"].ctypes.data if want_ids else None
        F[i].ao_open = o["ao_open"].ctypes.data if want_ao else None
        F[i].memory = 0
    bo = capi.BatchOptions()
    bo.ao_waves_per_simd = int(ao_waves_per_simd)
    bo.no_graph = 0 if graph else 1
    capi.check(lib.ubg_render_batch(device.handle, n, H, P, F, C.byref(bo)))
    del keeps
    res = []
    for i, o in enumerate(outs):
        fr = F[i]
        res.append(FrameResult(width=o["w"], height=o["h"], color=o["color"], depth=o["depth"], view_z=o["view_z"],
                               hit_id=o["hit_id"], ao_open=o["ao_open"], stats=fr.stats.as_dict(),
                               render_seconds=fr.render_seconds, post_seconds=fr.post_seconds,
                               cancelled=bool(fr.cancelled), layers=int(fr.layers),
                               kernel_ms={k: 0.0 for k in capi.KERNELS}, kernel_launches={k: 0 for k in capi.KERNELS},
                               aa_mask=None, aa_flagged=0, indirect=None, gi_occlusion=None, gi_features=None))
    return res


def host_query(scene: Scene, rays: np.ndarray, layout: int = capi.BVH8):
    """The library's traversal run on the CPU (test hook, no GPU needed):
    returns (closest hits, occluded flags, stats dict)."""
    lib = capi.load()
    desc, keep = scene.to_desc()
    rays = np.ascontiguousarray(rays, dtype=capi.ray_dtype())
    hits = np.zeros(len(rays), dtype=capi.hit_dtype())
    occ = np.zeros(len(rays), dtype=np.uint8)
    st = np.zeros(6, np.uint64)
    capi.check(lib.ubgx_host_query(C.byref(desc), int(layout), len(rays), rays.ctypes.data, hits.ctypes.data,
                                   occ.ctypes.data, st.ctypes.data))
    del keep
    names = ("node_visits", "prim_tests", "depth", "nodes", "stack_depth", "node_bytes")
    return hits, occ, {k: int(v) for k, v in zip(names, st)}


_default_device: Optional[Device] = None


def default_device() -> Device:
    global _default_device
    if _default_device is None:
        _default_device = Device(0)
    return _default_device


def render(scene: Scene, opt: RenderOptions, **kw) -> FrameResult:
    """umbreon::render(scene, opt): build, trace, resolve; one call."""
    dev = default_device()
    bs = BuiltScene(dev, scene)
    try:
        return bs.render(scene, opt, **kw)
    finally:
        bs.close()
