"""Object-space silhouette edges: umbreon::generateObjectSpaceEdges
(src/umbreon/edges/object_space_edges.cpp:172-264) over the C-ABI's
``ubg_object_space_edges``.

The generated edges are open cylinders with the flat outline material,
appended to a copy of the scene exactly as the reference appends them to
Scene::cylinders before tracing (render/pipeline.cpp:30-42).
"""
from __future__ import annotations

import ctypes as C
import copy
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import capi
from .scene import Material, Scene


@dataclass
class ObjectSpaceEdgeOptions:
    """umbreon::ObjectSpaceEdgeOptions (render/edge_types.hpp:311-409)."""
    enable: bool = False
    width: float = 0.03
    raise_: float = 0.0
    segments: int = 48
    color: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    clip: bool = True
    mesh_silhouette: bool = True
    mesh_crease: bool = True
    mesh_border: bool = True
    visibility_clip: bool = False
    mesh_hard_edge_deg: float = 40.0
    crease_angle_deg: float = 30.0
    mesh_crease_smooth_veto_deg: float = 0.0
    mesh_crease_convex_only: bool = False
    mesh_border_coplanar_veto_deg: float = 0.0
    mesh_crease_max_degree: int = 0

    def to_c(self) -> capi.EdgeOptions:
        o = capi.EdgeOptions()
        o.width, o.raise_, o.segments = self.width, self.raise_, int(self.segments)
        o.color[:] = [float(c) for c in self.color]
        o.clip = int(bool(self.clip))
        o.mesh_silhouette, o.mesh_crease = int(bool(self.mesh_silhouette)), int(bool(self.mesh_crease))
        o.mesh_border, o.visibility_clip = int(bool(self.mesh_border)), int(bool(self.visibility_clip))
        o.mesh_hard_edge_deg, o.crease_angle_deg = self.mesh_hard_edge_deg, self.crease_angle_deg
        o.mesh_crease_smooth_veto_deg = self.mesh_crease_smooth_veto_deg
        o.mesh_crease_convex_only = int(bool(self.mesh_crease_convex_only))
        o.mesh_border_coplanar_veto_deg = self.mesh_border_coplanar_veto_deg
        o.mesh_crease_max_degree = int(self.mesh_crease_max_degree)
        return o


def camera_c(scene: Scene) -> capi.Camera:
    c = scene.camera
    k = capi.Camera()
    k.position[:] = list(c.position)
    k.direction[:] = list(c.direction)
    k.up[:] = list(c.up)
    k.fovy, k.orthographic, k.height = c.fovy, int(bool(c.orthographic)), c.height
    return k


def edges_to_numpy(e: capi.Edges) -> dict:
    n = int(e.n)

    def arr(p, count, dt):
        if n == 0:
            return np.zeros((0,), dt)
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dt, copy=True)

    return {"p0": arr(e.p0, 3 * n, np.float32).reshape(n, 3), "p1": arr(e.p1, 3 * n, np.float32).reshape(n, 3),
            "color": arr(e.color, 4 * n, np.float32).reshape(n, 4), "opacity1": arr(e.opacity1, n, np.float32),
            "group": arr(e.group, n, np.uint16), "radius": float(e.radius)}


def object_space_edges(scene: Scene, opt: ObjectSpaceEdgeOptions, device=None) -> dict:
    """The generated edges as arrays (p0, p1, color, opacity1, group,
    radius). `device` (umbreon_amd.render.Device) runs the union clip; it
    may be None only when no clip runs."""
    lib = capi.load()
    desc, keep = scene.to_desc()
    cam, o = camera_c(scene), opt.to_c()
    out = capi.Edges()
    ctx = device.handle if device is not None else None
    capi.check(lib.ubg_object_space_edges(ctx, C.byref(desc), C.byref(cam), C.byref(o), C.byref(out)))
    try:
        return edges_to_numpy(out)
    finally:
        lib.ubg_edges_free(C.byref(out))
        del keep


def append_edges(scene: Scene, e: dict) -> Scene:
    """A copy of `scene` with the edges appended to its cylinders (open, flat
    outline material, not fromEdgeMacro)."""
    out = copy.deepcopy(scene)
    n = len(e["group"])
    if n == 0:
        return out
    m = out.material_index(Material.flat_outline())
    out.cyl_p0 = np.concatenate([out.cyl_p0, e["p0"]])
    out.cyl_p1 = np.concatenate([out.cyl_p1, e["p1"]])
    out.cyl_radius = np.concatenate([out.cyl_radius, np.full(n, e["radius"], np.float32)])
    out.cyl_color = np.concatenate([out.cyl_color, e["color"]])
    out.cyl_opacity1 = np.concatenate([out.cyl_opacity1, e["opacity1"]])
    out.cyl_material = np.concatenate([out.cyl_material, np.full(n, m, np.int32)])
    out.cyl_group = np.concatenate([out.cyl_group, e["group"]])
    out.cyl_open = np.concatenate([out.cyl_open, np.ones(n, np.uint8)])
    out.cyl_from_edge = np.concatenate([out.cyl_from_edge, np.zeros(n, np.uint8)])
    return out


def generate_object_space_edges(scene: Scene, opt: ObjectSpaceEdgeOptions, device=None) -> Scene:
    """generateObjectSpaceEdges on a copy: the scene with its edges appended
    (the scene itself when opt.enable is false)."""
    if not opt.enable:
        return scene
    return append_edges(scene, object_space_edges(scene, opt, device))
