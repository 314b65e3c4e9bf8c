"""Host side of the physically based path tracer (ubg_pathtrace).

No reference counterpart: the reference's integrators (pt1 / pt2) are
diffuse-only and refraction is outside its scope (docs/api/libumbreon.md:33).
The scene geometry is the ordinary umbreon Scene; the path tracer adds, per
material index, a surface model (Lambert, GGX conductor with Schlick Fresnel
of the primitive colour, GGX dielectric with exact Fresnel and refraction),
spherical area lights reached by next-event estimation (RIS over M
candidates, ReSTIR spatial reuse over K neighbours at the first vertex) and
a constant environment.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import capi
from .scene import Camera, Scene

DIFFUSE, CONDUCTOR, DIELECTRIC = capi.PT_DIFFUSE, capi.PT_CONDUCTOR, capi.PT_DIELECTRIC


@dataclass
class PtMaterial:
    type: int = DIFFUSE
    roughness: float = 0.5      # GGX alpha = roughness^2; < 0.02 is a smooth interface
    ior: float = 1.5


@dataclass
class PtOptions:
    width: int = 1920
    height: int = 1080
    spp: int = 1
    bounces: int = 4
    seed: int = 0
    # spherical lights: rows of (cx, cy, cz, radius, Lr, Lg, Lb)
    lights: np.ndarray = field(default_factory=lambda: np.zeros((0, 7), np.float32))
    env: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    prim_materials: List[PtMaterial] = field(default_factory=list)
    mesh_materials: List[PtMaterial] = field(default_factory=list)
    restir_candidates: int = 8
    restir_spatial: int = 4
    restir_radius: float = 16.0
    nee_all_lights: bool = False
    # network passes averaged by the learned denoiser (ubg_pt_params::
    # denoise_ensemble): 1, 2, 4, 8, or 0 = 8 from 32 spp on, else 1
    denoise_ensemble: int = 0


@dataclass
class PtResult:
    color: np.ndarray                 # H x W x 4 linear HDR
    features: Optional[np.ndarray]    # H x W x 12 or None
    trace_seconds: float
    denoise_seconds: float
    rays: int


def pt_params(camera: Camera, opt: PtOptions, denoise_net=None):
    """ubg_pt_params of the options (plus the arrays it points into)."""
    p = capi.PtParams()
    keep = []
    p.width, p.height = int(opt.width), int(opt.height)
    p.camera.position[:] = camera.position
    p.camera.direction[:] = camera.direction
    p.camera.up[:] = camera.up
    p.camera.fovy = camera.fovy
    p.camera.orthographic = int(bool(camera.orthographic))
    p.camera.height = camera.height
    p.spp, p.bounces, p.seed = int(opt.spp), int(opt.bounces), int(opt.seed) & 0xFFFFFFFF
    L = np.asarray(opt.lights, np.float32).reshape(-1, 7)
    lights = (capi.SphereLight * max(1, len(L)))()
    for i, row in enumerate(L):
        lights[i].center[:] = [float(v) for v in row[:3]]
        lights[i].radius = float(row[3])
        lights[i].radiance[:] = [float(v) for v in row[4:7]]
    keep.append(lights)
    p.lights = C.cast(lights, C.POINTER(capi.SphereLight))
    p.n_lights = len(L)
    p.env_radiance[:] = [float(v) for v in opt.env]

    def mats(ms: Sequence[PtMaterial]):
        arr = (capi.PtMaterial * max(1, len(ms)))()
        for i, m in enumerate(ms):
            arr[i].type, arr[i].roughness, arr[i].ior = int(m.type), float(m.roughness), float(m.ior)
        keep.append(arr)
        return C.cast(arr, C.POINTER(capi.PtMaterial)), len(ms)

    p.prim_materials, p.n_prim_materials = mats(opt.prim_materials)
    p.mesh_materials, p.n_mesh_materials = mats(opt.mesh_materials)
    p.restir_candidates = int(opt.restir_candidates)
    p.restir_spatial = int(opt.restir_spatial)
    p.restir_radius = float(opt.restir_radius)
    p.nee_all_lights = int(bool(opt.nee_all_lights))
    p.denoise_net = denoise_net.handle if denoise_net is not None else None
    p.denoise_ensemble = int(opt.denoise_ensemble)
    return p, keep


def pathtrace(built, scene: Scene, opt: PtOptions, want_features: bool = False, denoise_net=None) -> PtResult:
    """One path-traced image of a built scene (umbreon_amd.render.BuiltScene)
    with the scene's camera. denoise_net (an umbreon_amd.denoise.Unet whose
    input is the 12 features) replaces the colour by its denoising."""
    lib = capi.load()
    p, keep = pt_params(scene.camera, opt, denoise_net)
    color = np.empty((opt.height, opt.width, 4), np.float32)
    feats = np.empty((opt.height, opt.width, 12), np.float32) if want_features else None
    fr = capi.PtFrame()
    fr.color = color.ctypes.data
    fr.features = feats.ctypes.data if feats is not None else None
    fr.memory = 0
    capi.check(lib.ubg_pathtrace(built.device.handle, built.handle, C.byref(p), C.byref(fr)))
    del keep
    return PtResult(color=color, features=feats, trace_seconds=fr.trace_seconds,
                    denoise_seconds=fr.denoise_seconds, rays=int(fr.rays))
