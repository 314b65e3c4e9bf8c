"""Python mirror of umbreon's public scene types (src/umbreon/scene.hpp,
src/umbreon/render/render_options.hpp), stored structure-of-arrays.

Field names follow the reference (camelCase in C++, snake_case here). Only
the per-ray subset of RenderOptions is carried; options outside the gfx950
path (GI, adaptive AA, edges, denoise, enhanced AO) raise in :func:`params`
so a caller never gets a silently different image.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import capi

Vec3 = Tuple[float, float, float]


@dataclass
class Material:
    """umbreon::Material (scene.hpp:130-214): the POV finish fields, the
    shading model and the principled subset Material::Pbr (pbr_*)."""
    ambient: float = 0.2
    diffuse: float = 0.8
    specular: float = 0.0
    roughness: float = 0.02
    brilliance: float = 1.0
    phong: float = 0.0
    phong_size: float = 40.0
    metallic: bool = False
    reflection: float = 0.0
    emission: float = 0.0
    model: int = 0  # 0 Pov, 1 Principled, 2 Toon
    pbr_metallic: float = 0.0
    pbr_roughness: float = 0.5
    pbr_specular: float = 0.5
    pbr_anisotropy: float = 0.0
    pbr_anisotropy_rotation: float = 0.0  # turns

    @staticmethod
    def flat_outline() -> "Material":
        """Material::flatOutline(): ambient 1, diffuse 0 -> raw pigment."""
        return Material(ambient=1.0, diffuse=0.0, specular=0.0, roughness=0.02, brilliance=1.0,
                        phong=0.0, phong_size=40.0, metallic=False, reflection=0.0, emission=0.0)

    def to_c(self) -> capi.Material:
        m = capi.Material()
        m.ambient, m.diffuse, m.specular = self.ambient, self.diffuse, self.specular
        m.roughness, m.brilliance, m.phong = self.roughness, self.brilliance, self.phong
        m.phong_size, m.metallic = self.phong_size, int(bool(self.metallic))
        m.reflection, m.emission, m.model = self.reflection, self.emission, int(self.model)
        m.pbr_metallic, m.pbr_roughness, m.pbr_specular = self.pbr_metallic, self.pbr_roughness, self.pbr_specular
        m.pbr_anisotropy, m.pbr_anisotropy_rotation = self.pbr_anisotropy, self.pbr_anisotropy_rotation
        return m


@dataclass
class Camera:
    """umbreon::Camera (scene.hpp:380-388)."""
    position: Vec3 = (0.0, 0.0, 1.0)
    direction: Vec3 = (0.0, 0.0, -1.0)
    up: Vec3 = (0.0, 1.0, 0.0)
    fovy: float = 40.0
    orthographic: bool = False
    height: float = 1.0


@dataclass
class DistantLight:
    """umbreon::DistantLight (scene.hpp:390-411)."""
    direction: Vec3 = (0.0, -0.3, -1.0)
    color: Vec3 = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    casts_highlight: bool = True
    angular_radius: float = 0.0


@dataclass
class Fog:
    """umbreon::Fog (scene.hpp:413-428): OpenGL linear fog over view-z."""
    enabled: bool = False
    color: Vec3 = (0.0, 0.0, 0.0)
    start: float = 0.0
    end: float = 0.0


@dataclass
class RenderOptions:
    """Per-ray subset of umbreon::RenderOptions (render_options.hpp:14-328)."""
    width: int = 1024
    height: int = 768
    supersample: int = 1
    ao_samples: int = 0
    ao_distance: float = 1.0e20
    ao_intensity: float = 1.0
    shadows: bool = False
    shadow_samples: int = 1
    light_radius: float = 0.0  # degrees
    gi_integrator: int = 2
    specular_scale: float = 1.0
    transparency: bool = True
    transparent_background: bool = False
    max_transparent_layers: int = 256
    # enhanced AO (render_options.hpp:52-61); all defaults = legacy binary AO
    ao_falloff_power: float = 0.0
    ao_multi_scale: bool = False
    ao_bent_normal: bool = False
    ao_sky_color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    ao_ground_color: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    ao_use_camera_up: bool = True
    ao_up: Tuple[float, float, float] = (0.0, 1.0, 0.0)
    ao_multibounce: bool = False
    ao_low_discrepancy: bool = False
    ao_diffuse_factor: float = 0.0
    ao_res_div: int = 0  # coarse-grid AO: > 1 divisor, -1 = supersample factor, 0/1 = inline
    # adaptive antialiasing (render_options.hpp:20-39): 0 grid, 1 adaptive
    aa_mode: int = 0
    aa_threshold: float = 0.1
    aa_depth: int = 0
    aa_debug: bool = False  # fill FrameResult.aa_mask
    # path-traced GI (render_options.hpp:93-253); the gfx950 path implements
    # the pt1 and pt2 integrators (pt2 without ReSTIR, adaptive spp, emissive
    # NEE or traced reflections; E denoise = the a-trous filter) and refuses
    # the irradiance cache
    gi: bool = False
    gi_bounces: int = 1
    gi_max_distance: float = 0.0
    gi_intensity: float = 1.0
    gi_env_intensity: float = 1.0
    pt1_spp: int = 8
    pt1_seed: int = 0
    pt1_sky_mode: int = 0
    pt1_sky_radiance: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    pt1_ld: bool = True
    pt1_clamp: float = 0.0
    pt1_gather_div: int = -1
    pt1_half_res: bool = True
    pt1_denoise: bool = True
    denoise_iters: int = 5
    denoise_sigma_z: float = 1.0
    denoise_sigma_n: float = 128.0
    denoise_sigma_l: float = 4.0
    pt1_upsample_normal_pow: float = 32.0
    pt1_upsample_depth_scale: float = 0.02
    pt1_edge_patch: bool = True
    pt1_edge_patch_spp_mul: int = 4
    pt1_edge_patch_thresh: float = 0.3
    pt2_pattern: int = 1
    pt2_emissive: bool = True
    pt2_rounds: int = 0
    pt2_adaptive: bool = False
    pt2_emissive_nee: bool = True
    pt2_reflect: bool = True
    # options the gfx950 path does not implement; must stay at their defaults
    denoiser: int = 0
    env_lights: int = 0


def _f32(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape_last is not None and a.size:
        a = a.reshape(-1, shape_last)
    return a


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None or a.size == 0:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Scene:
    """umbreon::Scene in structure-of-arrays form."""
    # Mesh
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    normals: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    colors: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    index: Optional[np.ndarray] = None          # (T, 3) uint32 or None (soup)
    mesh_material: Material = field(default_factory=Material)
    mesh_materials: List[Material] = field(default_factory=list)
    tri_material: Optional[np.ndarray] = None   # (T,) uint8
    tri_group: Optional[np.ndarray] = None      # (T,) uint16
    # Spheres
    sphere_center: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    sphere_radius: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))
    sphere_color: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    sphere_material: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    sphere_group: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint16))
    sphere_from_edge: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint8))
    # Cylinders
    cyl_p0: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    cyl_p1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    cyl_radius: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))
    cyl_color: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    cyl_opacity1: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.float32))
    cyl_material: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    cyl_group: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint16))
    cyl_open: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint8))
    cyl_from_edge: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.uint8))
    prim_materials: List[Material] = field(default_factory=list)
    instance_offsets: Optional[np.ndarray] = None
    # Frame-level
    camera: Camera = field(default_factory=Camera)
    lights: List[DistantLight] = field(default_factory=list)
    ambient_color: Vec3 = (1.0, 1.0, 1.0)
    background: Vec3 = (0.0, 0.0, 0.0)
    fog: Fog = field(default_factory=Fog)
    assumed_gamma: float = 1.0
    clip_near: float = -math.inf
    clip_far: float = math.inf

    # -- building helpers -------------------------------------------------
    def material_index(self, m: Material) -> int:
        for i, x in enumerate(self.prim_materials):
            if x == m:
                return i
        self.prim_materials.append(m)
        return len(self.prim_materials) - 1

    def add_sphere(self, center, radius, color, material: Optional[Material] = None,
                   group: int = 0, from_edge: bool = False) -> None:
        self.add_spheres([center], [radius], [color], material, group, from_edge)

    def add_spheres(self, centers, radii, colors, material: Optional[Material] = None,
                    group=0, from_edge=False) -> None:
        m = self.material_index(material if material is not None else Material.flat_outline())
        c = _f32(centers, 3)
        n = len(c)
        self.sphere_center = np.concatenate([self.sphere_center, c])
        self.sphere_radius = np.concatenate([self.sphere_radius, _f32(radii).reshape(n)])
        self.sphere_color = np.concatenate([self.sphere_color, _f32(colors, 4)])
        self.sphere_material = np.concatenate([self.sphere_material, np.full(n, m, np.int32)])
        self.sphere_group = np.concatenate([self.sphere_group, np.broadcast_to(np.asarray(group, np.uint16), (n,))])
        self.sphere_from_edge = np.concatenate(
            [self.sphere_from_edge, np.broadcast_to(np.asarray(from_edge, np.uint8), (n,))])

    def add_cylinder(self, p0, p1, radius, color, material: Optional[Material] = None,
                     group: int = 0, opacity1: float = -1.0, open: bool = False,
                     from_edge: bool = False) -> None:
        self.add_cylinders([p0], [p1], [radius], [color], material, group, [opacity1], open, from_edge)

    def add_cylinders(self, p0, p1, radii, colors, material: Optional[Material] = None, group=0,
                      opacity1=None, open=False, from_edge=False) -> None:
        m = self.material_index(material if material is not None else Material.flat_outline())
        a = _f32(p0, 3)
        n = len(a)
        op1 = np.full(n, -1.0, np.float32) if opacity1 is None else _f32(opacity1).reshape(n)
        self.cyl_p0 = np.concatenate([self.cyl_p0, a])
        self.cyl_p1 = np.concatenate([self.cyl_p1, _f32(p1, 3)])
        self.cyl_radius = np.concatenate([self.cyl_radius, _f32(radii).reshape(n)])
        self.cyl_color = np.concatenate([self.cyl_color, _f32(colors, 4)])
        self.cyl_opacity1 = np.concatenate([self.cyl_opacity1, op1])
        self.cyl_material = np.concatenate([self.cyl_material, np.full(n, m, np.int32)])
        self.cyl_group = np.concatenate([self.cyl_group, np.broadcast_to(np.asarray(group, np.uint16), (n,))])
        self.cyl_open = np.concatenate([self.cyl_open, np.broadcast_to(np.asarray(open, np.uint8), (n,))])
        self.cyl_from_edge = np.concatenate(
            [self.cyl_from_edge, np.broadcast_to(np.asarray(from_edge, np.uint8), (n,))])

    def set_mesh(self, positions, normals, colors, index=None, material: Optional[Material] = None,
                 materials: Optional[Sequence[Material]] = None, tri_material=None,
                 tri_group=None) -> None:
        self.positions = _f32(positions, 3)
        self.normals = _f32(normals, 3)
        self.colors = _f32(colors, 4)
        self.index = None if index is None else np.ascontiguousarray(np.asarray(index, np.uint32).reshape(-1, 3))
        self.mesh_material = material if material is not None else Material()
        self.mesh_materials = list(materials) if materials else []
        self.tri_material = None if tri_material is None else np.ascontiguousarray(np.asarray(tri_material, np.uint8))
        self.tri_group = None if tri_group is None else np.ascontiguousarray(np.asarray(tri_group, np.uint16))

    @property
    def n_triangles(self) -> int:
        if self.index is not None:
            return int(len(self.index))
        return int(len(self.positions) // 3)

    # -- C-ABI marshalling ------------------------------------------------
    def to_desc(self) -> Tuple[capi.SceneDesc, list]:
        keep: list = []
        d = capi.SceneDesc()

        def arr(a, dt):
            a = np.ascontiguousarray(np.asarray(a, dtype=dt))
            keep.append(a)
            return a

        pos = arr(self.positions, np.float32)
        d.n_vertices = len(pos)
        d.positions = _ptr(pos, C.c_float)
        nrm = arr(self.normals, np.float32) if len(self.normals) == len(pos) else arr(np.zeros_like(pos), np.float32)
        d.normals = _ptr(nrm, C.c_float)
        col = arr(self.colors, np.float32) if len(self.colors) == len(pos) else arr(
            np.tile(np.array([0, 0, 0, 1], np.float32), (len(pos), 1)), np.float32)
        d.colors = _ptr(col, C.c_float)
        d.n_triangles = self.n_triangles if len(pos) else 0
        if self.index is not None:
            idx = arr(self.index, np.uint32)
            d.index = _ptr(idx, C.c_uint32)
        else:
            d.index = C.cast(None, C.POINTER(C.c_uint32))
        mats = self.mesh_materials if self.tri_material is not None else [self.mesh_material]
        mm = (capi.Material * max(1, len(mats)))(*[m.to_c() for m in mats])
        keep.append(mm)
        d.mesh_materials = C.cast(mm, C.POINTER(capi.Material))
        d.n_mesh_materials = len(mats)
        d.tri_material = _ptr(arr(self.tri_material, np.uint8), C.c_uint8) if self.tri_material is not None else C.cast(None, u8p())
        d.tri_group = _ptr(arr(self.tri_group, np.uint16), C.c_uint16) if self.tri_group is not None else C.cast(None, u16p())
        n = len(self.sphere_center)
        d.n_spheres = n
        if n:
            cr = arr(np.concatenate([self.sphere_center, self.sphere_radius.reshape(-1, 1)], axis=1), np.float32)
            d.sphere_center_radius = _ptr(cr, C.c_float)
            d.sphere_color = _ptr(arr(self.sphere_color, np.float32), C.c_float)
            d.sphere_material = _ptr(arr(self.sphere_material, np.int32), C.c_int32)
            d.sphere_group = _ptr(arr(self.sphere_group, np.uint16), C.c_uint16)
            d.sphere_from_edge = _ptr(arr(self.sphere_from_edge, np.uint8), C.c_uint8)
        n = len(self.cyl_p0)
        d.n_cylinders = n
        if n:
            d.cyl_p0 = _ptr(arr(self.cyl_p0, np.float32), C.c_float)
            d.cyl_p1 = _ptr(arr(self.cyl_p1, np.float32), C.c_float)
            d.cyl_radius = _ptr(arr(self.cyl_radius, np.float32), C.c_float)
            d.cyl_color = _ptr(arr(self.cyl_color, np.float32), C.c_float)
            d.cyl_opacity1 = _ptr(arr(self.cyl_opacity1, np.float32), C.c_float)
            d.cyl_material = _ptr(arr(self.cyl_material, np.int32), C.c_int32)
            d.cyl_group = _ptr(arr(self.cyl_group, np.uint16), C.c_uint16)
            d.cyl_open = _ptr(arr(self.cyl_open, np.uint8), C.c_uint8)
            d.cyl_from_edge = _ptr(arr(self.cyl_from_edge, np.uint8), C.c_uint8)
        pm = (capi.Material * max(1, len(self.prim_materials)))(*[m.to_c() for m in self.prim_materials])
        keep.append(pm)
        d.prim_materials = C.cast(pm, C.POINTER(capi.Material))
        d.n_prim_materials = len(self.prim_materials)
        if self.instance_offsets is not None and len(self.instance_offsets):
            io = arr(self.instance_offsets, np.float32)
            d.n_instances = len(io)
            d.instance_offsets = _ptr(io, C.c_float)
        return d, keep


def u8p():
    return C.POINTER(C.c_uint8)


def u16p():
    return C.POINTER(C.c_uint16)


def params(scene: Scene, opt: RenderOptions, flags: int = 0,
           cancel_flag: Optional[C.c_int32] = None) -> Tuple[capi.RenderParams, list]:
    """RenderOptions + Scene frame constants -> ubg_render_params."""
    if opt.denoiser != 0 or opt.env_lights > 0:
        raise capi.UmbreonError(capi.UBG_ERR_UNSUPPORTED,
                                "the final-colour denoiser / env lights are not part of the gfx950 path")
    keep: list = []
    p = capi.RenderParams()
    p.width, p.height, p.supersample = int(opt.width), int(opt.height), int(opt.supersample)
    cam = scene.camera
    p.camera.position[:] = cam.position
    p.camera.direction[:] = cam.direction
    p.camera.up[:] = cam.up
    p.camera.fovy = cam.fovy
    p.camera.orthographic = int(bool(cam.orthographic))
    p.camera.height = cam.height
    lights = (capi.Light * max(1, len(scene.lights)))()
    for i, l in enumerate(scene.lights):
        lights[i].direction[:] = l.direction
        lights[i].color[:] = l.color
        lights[i].intensity = l.intensity
        lights[i].casts_highlight = int(bool(l.casts_highlight))
        lights[i].angular_radius = l.angular_radius
    keep.append(lights)
    p.lights = C.cast(lights, C.POINTER(capi.Light))
    p.n_lights = len(scene.lights)
    p.ambient_color[:] = scene.ambient_color
    p.background[:] = scene.background
    p.fog.enabled = int(bool(scene.fog.enabled))
    p.fog.color[:] = scene.fog.color
    p.fog.start, p.fog.end = scene.fog.start, scene.fog.end
    p.clip_near, p.clip_far = scene.clip_near, scene.clip_far
    p.assumed_gamma = scene.assumed_gamma
    p.ao_samples, p.ao_distance, p.ao_intensity = int(opt.ao_samples), opt.ao_distance, opt.ao_intensity
    p.shadows, p.shadow_samples = int(bool(opt.shadows)), int(opt.shadow_samples)
    p.light_radius_deg, p.gi_integrator = opt.light_radius, int(opt.gi_integrator)
    p.specular_scale = opt.specular_scale
    p.transparency = int(bool(opt.transparency))
    p.transparent_background = int(bool(opt.transparent_background))
    p.max_transparent_layers = int(opt.max_transparent_layers)
    p.ao_falloff_power = opt.ao_falloff_power
    p.ao_multi_scale = int(bool(opt.ao_multi_scale))
    p.ao_bent_normal = int(bool(opt.ao_bent_normal))
    p.ao_sky_color[:] = opt.ao_sky_color
    p.ao_ground_color[:] = opt.ao_ground_color
    p.ao_use_camera_up = int(bool(opt.ao_use_camera_up))
    p.ao_up[:] = opt.ao_up
    p.ao_multibounce = int(bool(opt.ao_multibounce))
    p.ao_low_discrepancy = int(bool(opt.ao_low_discrepancy))
    p.ao_diffuse_factor = opt.ao_diffuse_factor
    p.ao_res_div = int(opt.ao_res_div)
    p.aa_mode = int(opt.aa_mode)
    p.aa_threshold = float(opt.aa_threshold)
    p.aa_depth = int(opt.aa_depth)
    p.gi = int(bool(opt.gi))
    p.gi_bounces = int(opt.gi_bounces)
    p.gi_max_distance = float(opt.gi_max_distance)
    p.gi_intensity = float(opt.gi_intensity)
    p.gi_env_intensity = float(opt.gi_env_intensity)
    p.pt1_spp = int(opt.pt1_spp)
    p.pt1_seed = int(opt.pt1_seed) & 0xFFFFFFFF
    p.pt1_sky_mode = int(opt.pt1_sky_mode)
    p.pt1_sky_radiance[:] = opt.pt1_sky_radiance
    p.pt1_ld = int(bool(opt.pt1_ld))
    p.pt1_clamp = float(opt.pt1_clamp)
    p.pt1_gather_div = int(opt.pt1_gather_div)
    p.pt1_half_res = int(bool(opt.pt1_half_res))
    p.pt1_denoise = int(bool(opt.pt1_denoise))
    p.denoise_iters = int(opt.denoise_iters)
    p.denoise_sigma_z = float(opt.denoise_sigma_z)
    p.denoise_sigma_n = float(opt.denoise_sigma_n)
    p.denoise_sigma_l = float(opt.denoise_sigma_l)
    p.pt1_upsample_normal_pow = float(opt.pt1_upsample_normal_pow)
    p.pt1_upsample_depth_scale = float(opt.pt1_upsample_depth_scale)
    p.pt1_edge_patch = int(bool(opt.pt1_edge_patch))
    p.pt1_edge_patch_spp_mul = int(opt.pt1_edge_patch_spp_mul)
    p.pt1_edge_patch_thresh = float(opt.pt1_edge_patch_thresh)
    p.pt2_pattern = int(opt.pt2_pattern)
    p.pt2_emissive = int(bool(opt.pt2_emissive))
    p.pt2_rounds = int(opt.pt2_rounds)
    p.pt2_adaptive = int(bool(opt.pt2_adaptive))
    p.pt2_emissive_nee = int(bool(opt.pt2_emissive_nee))
    p.pt2_reflect = int(bool(opt.pt2_reflect))
    p.flags = flags
    if cancel_flag is not None:
        keep.append(cancel_flag)
        p.cancel_flag = C.pointer(cancel_flag)
    return p, keep
