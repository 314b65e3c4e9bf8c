"""Known-answer cases restated from the reference's own test suite.

Each case builds one or more small synthetic scenes (the reference's tests use
5x5 ortho views of quads and spheres with closed-form centre pixels) and a
check over the rendered frames. The same cases run against the CPU oracle
(tests/test_oracle.py) and the gfx950 path (tests/test_gpu_refcases.py).

Fixture builders follow tests/render_test_util.hpp:14-71 of the reference;
every case cites the assertion it restates.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, Tuple

import numpy as np

from umbreon_amd.scene import Camera, DistantLight, Fog, Material, RenderOptions, Scene

PIGMENT = (0.5, 0.6, 0.7, 1.0)
C = (2, 2)  # centre pixel (row, col) of a 5x5 frame


def quad_mesh(color, z=0.0):
    """makeQuad (render_test_util.hpp:14-28): [-2,2]^2 at z, +Z, 2 tris."""
    p = [(-2, -2, z), (2, -2, z), (2, 2, z), (-2, -2, z), (2, 2, z), (-2, 2, z)]
    return p, [(0, 0, 1)] * 6, [tuple(color)] * 6


def ortho_cam(z=10.0) -> Camera:
    """makeOrthoCam (render_test_util.hpp:32-40)."""
    return Camera(position=(0, 0, z), direction=(0, 0, -1), up=(0, 1, 0), orthographic=True, height=4.0)


def key_light() -> DistantLight:
    """makeKeyLight (render_test_util.hpp:45-51): head-on, intensity 0.5."""
    return DistantLight(direction=(0, 0, -1), color=(1, 1, 1), intensity=0.5)


def quad_scene(color=PIGMENT, bg=(0.4, 0.1, 0.2)) -> Scene:
    sc = Scene()
    p, n, c = quad_mesh(color)
    sc.set_mesh(p, n, c, material=Material(ambient=0.2, diffuse=0.8))
    sc.camera = ortho_cam()
    sc.lights = [key_light()]
    sc.background = bg
    return sc


def material_sphere_scene(color, mat: Material, bg) -> Scene:
    """makeMaterialSphereScene (render_test_util.hpp:57-71)."""
    sc = Scene()
    sc.camera = ortho_cam()
    sc.lights = [key_light()]
    sc.background = bg
    sc.add_sphere((0, 0, 0), 1.0, color, mat)
    return sc


def opts(w=5, h=5, **kw) -> RenderOptions:
    return RenderOptions(width=w, height=h, **kw)


@dataclass
class Case:
    name: str
    renders: Dict[str, Tuple[Scene, RenderOptions]]
    check: Callable[[dict], None]
    # the GPU must match the oracle per pixel within this max abs error
    tol: float = 1e-4


def approx(a, b, eps):
    assert abs(float(a) - float(b)) <= eps, f"{a} != {b} (+-{eps})"


CASES = []


def case(fn):
    CASES.append(fn())
    return fn


# ---------------------------------------------------------------- shading
@case
def mesh_direct():
    # test_render_shading.cpp:19-36
    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.30, 1e-4)
        approx(f.color[C][1], 0.36, 1e-4)
        approx(f.color[C][2], 0.42, 1e-4)
        approx(f.color[C][3], 1.0, 1e-6)
        approx(f.depth[C], 10.0, 1e-3)
    return Case("mesh_direct", {"f": (quad_scene(), opts())}, chk)


@case
def background_miss():
    # test_render_shading.cpp:39-53
    sc = Scene()
    sc.camera = ortho_cam()
    sc.lights = [key_light()]
    sc.background = (0.4, 0.1, 0.2)

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.4, 1e-6)
        approx(f.color[C][1], 0.1, 1e-6)
        approx(f.color[C][2], 0.2, 1e-6)
        approx(f.depth[C], 0.0, 1e-6)
    return Case("background_miss", {"f": (sc, opts())}, chk)


@case
def outline_sphere():
    # test_render_shading.cpp:58-76 (default flatOutline material = raw colour)
    sc = Scene()
    sc.camera = ortho_cam()
    sc.lights = [key_light()]
    sc.background = (0.4, 0.1, 0.2)
    sc.add_sphere((0, 0, 0), 1.0, (0.3, 0.4, 0.5, 1.0))

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.3, 1e-4)
        approx(f.color[C][1], 0.4, 1e-4)
        approx(f.color[C][2], 0.5, 1e-4)
    return Case("outline_sphere", {"f": (sc, opts())}, chk)


@case
def assumed_gamma():
    # test_render_shading.cpp:80-94
    sc = quad_scene(bg=(0, 0, 0))
    sc.assumed_gamma = 2.0

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.30 * 0.30, 1e-4)
        approx(f.color[C][1], 0.36 * 0.36, 1e-4)
        approx(f.color[C][3], 1.0, 1e-6)
    return Case("assumed_gamma", {"f": (sc, opts())}, chk)


@case
def supersample_uniform():
    # test_render_shading.cpp:98-114
    def chk(r):
        f = r["f"]
        assert f.width == 5 and f.height == 5
        approx(f.color[C][0], 0.30, 1e-4)
    return Case("supersample_uniform", {"f": (quad_scene(bg=(0, 0, 0)), opts(supersample=2))}, chk)


@case
def matte_sphere():
    # test_render_shading.cpp:119-132
    m = Material(ambient=0.0, diffuse=0.8, brilliance=0.0, specular=0.0, phong=0.0)

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.24, 1e-4)
        approx(f.color[C][1], 0.16, 1e-4)
        approx(f.color[C][2], 0.08, 1e-4)
    return Case("matte_sphere", {"f": (material_sphere_scene((0.6, 0.4, 0.2, 1.0), m, (0, 0, 0)), opts())}, chk)


METAL = Material(ambient=0.35, diffuse=0.3, brilliance=2.0, specular=0.80, roughness=1.0 / 20.0,
                 metallic=True, phong=0.0, reflection=0.10)


@case
def metallic_sphere():
    # test_render_shading.cpp:137-149
    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.54, 1e-3)
        approx(f.color[C][1], 0.36, 1e-3)
        approx(f.color[C][2], 0.18, 1e-3)
    return Case("metallic_sphere", {"f": (material_sphere_scene((0.6, 0.4, 0.2, 1.0), METAL, (0, 0, 0)), opts())},
                chk)


@case
def reflection_term():
    # test_render_shading.cpp:153-164
    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.64, 1e-3)
        approx(f.color[C][1], 0.46, 1e-3)
    return Case("reflection_term", {"f": (material_sphere_scene((0.6, 0.4, 0.2, 1.0), METAL, (1, 1, 1)), opts())},
                chk)


@case
def phong_lift():
    # test_render_shading.cpp:171-186
    m = Material(ambient=0.3, diffuse=0.5, brilliance=0.0, specular=0.0, phong=10000.0, phong_size=50.0)

    def chk(r):
        f = r["f"]
        lr, lg, lb = f.color[C][0] - 0.11, f.color[C][1] - 0.22, f.color[C][2] - 0.33
        approx(lr, 5000.0, 1.0)
        approx(lr, lg, 1e-3)
        approx(lg, lb, 1e-3)
    return Case("phong_lift", {"f": (material_sphere_scene((0.2, 0.4, 0.6, 1.0), m, (0, 0, 0)), opts())}, chk,
                tol=2e-3)


@case
def fill_light():
    # test_render_shading.cpp:192-213
    m = Material(ambient=0.0, diffuse=0.0, brilliance=1.0, specular=0.5, roughness=0.05, phong=0.0)
    lit = material_sphere_scene((0.6, 0.4, 0.2, 1.0), m, (0, 0, 0))
    fill = material_sphere_scene((0.6, 0.4, 0.2, 1.0), m, (0, 0, 0))
    fill.lights[0].casts_highlight = False

    def chk(r):
        approx(r["lit"].color[C][0], 0.25, 1e-3)
        approx(r["fill"].color[C][0], 0.0, 1e-4)
        approx(r["fill"].color[C][1], 0.0, 1e-4)
    return Case("fill_light", {"lit": (lit, opts()), "fill": (fill, opts())}, chk)


@case
def metallic_fresnel():
    # test_render_shading.cpp:220-242
    sc = Scene()
    p, n, c = quad_mesh((1, 0, 0, 1))
    sc.set_mesh(p, n, c, material=Material(ambient=0.0, diffuse=0.0, specular=1.0, roughness=0.5,
                                            brilliance=1.0, metallic=True))
    sc.camera = ortho_cam()
    sc.lights = [DistantLight(direction=(-0.8660254, 0.0, -0.5), color=(1, 1, 1), intensity=1.0)]

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.75, 2e-3)
        approx(f.color[C][1], 0.04445, 2e-3)
        approx(f.color[C][1], f.color[C][2], 1e-5)
    return Case("metallic_fresnel", {"f": (sc, opts())}, chk)


# ---------------------------------------------------------------- shadows
def floor_slab_mesh(with_slab=True, mat=None):
    p = [(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, -2, 0), (2, 2, 0), (-2, 2, 0)]
    n = [(0, 0, 1)] * 6
    c = [(1, 1, 1, 1)] * 6
    if with_slab:
        z = 0.4
        p += [(0.1, -2, z), (4, -2, z), (4, 2, z), (0.1, -2, z), (4, 2, z), (0.1, 2, z)]
        n += [(0, 0, -1)] * 6
        c += [(0, 0, 0, 1)] * 6
    return p, n, c


@case
def hard_shadow():
    # test_render_shadows_gi.cpp:27-74
    sc = Scene()
    p, n, c = floor_slab_mesh(True)
    sc.set_mesh(p, n, c, material=Material(ambient=0.2, diffuse=0.8))
    sc.camera = ortho_cam()
    sc.lights = [DistantLight(direction=(-0.6, 0.0, -0.8), color=(1, 1, 1), intensity=1.0)]

    def chk(r):
        approx(r["off"].color[C][0], 0.84, 0.03)
        approx(r["on"].color[C][0], 0.2, 0.03)
        assert r["on"].color[C][0] + 0.2 < r["off"].color[C][0]
    return Case("hard_shadow", {"off": (sc, opts(shadows=False)), "on": (sc, opts(shadows=True))}, chk)


@case
def soft_shadow():
    # test_render_shadows_gi.cpp:80-117
    sc = quad_scene(color=(1, 1, 1, 1), bg=(0, 0, 0))
    sc.lights = [DistantLight(direction=(-0.6, 0.0, -0.8), color=(1, 1, 1), intensity=1.0)]
    sc.add_sphere((0.9, 0.0, 1.2), 0.45, (0, 0, 0, 1.0))

    def chk(r):
        lit, hard, soft = (r[k].color[C][0] for k in ("lit", "hard", "soft"))
        approx(hard, 0.2, 0.03)
        approx(lit, 0.84, 0.03)
        assert hard + 0.05 < soft and soft + 0.05 < lit
        assert np.array_equal(r["soft"].color, r["soft2"].color)
    soft = opts(shadows=True, light_radius=22.0, shadow_samples=64)
    return Case("soft_shadow", {"lit": (sc, opts(shadows=False)), "hard": (sc, opts(shadows=True)),
                                "soft": (sc, soft), "soft2": (sc, soft)}, chk, tol=1e-3)


@case
def acne_guard():
    # test_render_shadows_gi.cpp:129-158
    sc = Scene()
    q = 100.0
    p = [(-q, -q, 0), (q, -q, 0), (q, q, 0), (-q, -q, 0), (q, q, 0), (-q, q, 0)]
    sc.set_mesh(p, [(0, 0, 1)] * 6, [(1, 1, 1, 1)] * 6, material=Material(ambient=0.2, diffuse=0.8))
    sc.camera = Camera(position=(0.0, 1200.0, 1600.0), direction=(0.0, -0.6, -0.8), up=(0, 1, 0),
                       orthographic=True, height=6.0)
    sc.lights = [key_light()]

    def chk(r):
        assert r["f"].color[:, :, 0].min() > 0.5
    return Case("acne_guard", {"f": (sc, opts(shadows=True))}, chk)


# ---------------------------------------------------------------- AO
@case
def ao_open():
    # test_render_ao.cpp:28-41
    sc = quad_scene(bg=(0, 0, 0))

    def chk(r):
        approx(r["f"].color[C][0], 0.30, 0.02)
        approx(r["f"].color[C][1], 0.36, 0.02)
    return Case("ao_open", {"f": (sc, opts(ao_samples=64, ao_distance=100.0))}, chk)


@case
def ao_slab():
    # test_render_ao.cpp:47-93
    def scene(with_slab):
        sc = Scene()
        p, n, c = floor_slab_mesh(with_slab)
        sc.set_mesh(p, n, c, material=Material(ambient=1.0, diffuse=0.0))
        sc.camera = ortho_cam()
        return sc
    o = opts(ao_samples=128, ao_distance=10.0)

    def chk(r):
        approx(r["open"].color[C][0], 1.0, 0.03)
        assert r["occ"].color[C][0] + 0.05 < r["open"].color[C][0]
        assert r["occ"].color[C][0] > 0.05
    return Case("ao_slab", {"open": (scene(False), o), "occ": (scene(True), o)}, chk)


@case
def ao_outline_gate():
    # test_render_ao.cpp:103-148
    def scene(probed_from_edge):
        sc = Scene()
        sc.camera = ortho_cam()
        m = Material(ambient=1.0, diffuse=0.0)
        sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1), m, from_edge=probed_from_edge)
        sc.add_sphere((0, 2.2, 0), 2.0, (0, 0, 0, 1), m, from_edge=True)
        return sc
    off = opts(ao_samples=0)
    on = opts(ao_samples=128, ao_distance=10.0)

    def chk(r):
        assert np.abs(r["rOn"].color - r["rOff"].color).sum() > 0.05
        assert np.array_equal(r["oOn"].color, r["oOff"].color)
    return Case("ao_outline_gate", {"rOff": (scene(False), off), "rOn": (scene(False), on),
                                    "oOff": (scene(True), off), "oOn": (scene(True), on)}, chk)


@case
def ao_determinism():
    # test_render_ao.cpp:150-168
    sc = quad_scene(bg=(0, 0, 0))
    o = opts(ao_samples=32, ao_distance=100.0)

    def chk(r):
        assert np.array_equal(r["a"].color, r["b"].color)
    return Case("ao_determinism", {"a": (sc, o), "b": (sc, o)}, chk)


# ---------------------------------------------------------------- enhanced AO
def _ao_floor(slab_z=None, gray=1.0, mat=None):
    p = [(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, -2, 0), (2, 2, 0), (-2, 2, 0)]
    n = [(0, 0, 1)] * 6
    c = [(gray, gray, gray, 1)] * 6
    if slab_z is not None:
        z = slab_z
        p += [(0.1, -2, z), (4, -2, z), (4, 2, z), (0.1, -2, z), (4, 2, z), (0.1, 2, z)]
        n += [(0, 0, -1)] * 6
        c += [(0, 0, 0, 1)] * 6
    sc = Scene()
    sc.set_mesh(p, n, c, material=mat or Material(ambient=1.0, diffuse=0.0))
    sc.camera = ortho_cam()
    sc.ambient_color = (1, 1, 1)
    sc.background = (0, 0, 0)
    return sc


@case
def ao_falloff_multiscale():
    # test_render_ao.cpp:170-231 (distance falloff and multi-scale vs binary)
    def o(falloff, ms):
        return opts(ao_samples=256, ao_distance=10.0, ao_falloff_power=falloff, ao_multi_scale=ms)
    r = {"binNear": (_ao_floor(0.4), o(0.0, False)), "binFar": (_ao_floor(3.0), o(0.0, False)),
         "foNear": (_ao_floor(0.4), o(2.0, False)), "foFar": (_ao_floor(3.0), o(2.0, False)),
         "msFar": (_ao_floor(4.0), o(0.0, True)), "msNear": (_ao_floor(0.4), o(0.0, True)),
         "binFar4": (_ao_floor(4.0), o(0.0, False))}

    def chk(f):
        v = {k: float(f[k].color[C][0]) for k in f}
        assert v["foNear"] >= v["binNear"] - 1e-4 and v["foFar"] >= v["binFar"] - 1e-4, v
        assert (v["foFar"] - v["binFar"]) > (v["foNear"] - v["binNear"]) + 0.05, v
        assert v["msFar"] > v["binFar4"] + 0.05 and v["msNear"] < v["msFar"], v
    return Case("ao_falloff_multiscale", r, chk)


@case
def ao_bent_normal():
    # test_render_ao.cpp:233-278 (sky/ground gradient along the bent normal)
    def o(up):
        return opts(ao_samples=64, ao_distance=10.0, ao_bent_normal=True, ao_sky_color=(1, 1, 1),
                    ao_ground_color=(0, 0, 0), ao_use_camera_up=False, ao_up=up)
    r = {"plus": (_ao_floor(), o((0, 0, 1))), "minus": (_ao_floor(), o((0, 0, -1))),
         "perp": (_ao_floor(), o((1, 0, 0)))}

    def chk(f):
        up, um, pp = (float(f[k].color[C][0]) for k in ("plus", "minus", "perp"))
        assert up > 0.9 and um < 0.1, (up, um)
        approx(pp, 0.5, 0.1)
        assert um < pp < up
    return Case("ao_bent_normal", r, chk)


@case
def ao_multibounce():
    # test_render_ao.cpp:280-336 (albedo-aware lift)
    def o(mb):
        return opts(ao_samples=256, ao_distance=10.0, ao_multibounce=mb)
    r = {"openW": (_ao_floor(None, 1.0), o(True)), "occW": (_ao_floor(0.4, 1.0), o(True)),
         "openG": (_ao_floor(None, 0.5), o(True)), "occG": (_ao_floor(0.4, 0.5), o(True)),
         "occWplain": (_ao_floor(0.4, 1.0), o(False))}

    def chk(f):
        v = {k: float(f[k].color[C][0]) for k in f}
        assert v["occW"] / v["openW"] > v["occG"] / v["openG"] + 0.02, v
        assert v["occW"] > v["occWplain"] + 0.01, v
    return Case("ao_multibounce", r, chk)


@case
def ao_diffuse_factor():
    # test_render_ao.cpp:338-380 (aoDiffuseFactor darkens occluded direct diffuse)
    def scene():
        sc = _ao_floor(0.4, 1.0, Material(ambient=0.2, diffuse=0.8))
        sc.lights = [key_light()]
        return sc
    r = {"on": (scene(), opts(ao_samples=256, ao_distance=10.0, ao_diffuse_factor=0.7)),
         "off": (scene(), opts(ao_samples=256, ao_distance=10.0, ao_diffuse_factor=0.0))}

    def chk(f):
        assert f["on"].color[C][0] + 0.01 < f["off"].color[C][0]
    return Case("ao_diffuse_factor", r, chk)


@case
def ao_low_discrepancy():
    # test_render_ao.cpp:382-436 (Hammersley + per-pixel rotation beats white
    # noise against a converged reference; multi-scale keeps all on the
    # enhanced estimator)
    def o(n, ld):
        return opts(ao_samples=n, ao_distance=10.0, ao_multi_scale=True, ao_low_discrepancy=ld)
    sc = _ao_floor(0.6)
    r = {"ref": (sc, o(2048, False)), "noise": (sc, o(32, False)), "ld": (sc, o(32, True))}

    def chk(f):
        ref = f["ref"].color[..., 0]
        assert np.abs(f["ld"].color[..., 0] - ref).sum() < np.abs(f["noise"].color[..., 0] - ref).sum()
    return Case("ao_low_discrepancy", r, chk)


# ---------------------------------------------------------------- coarse-grid AO
def _coarse_scene(shell=False):
    """makeAoScene (test_ao_coarse.cpp:37-72): ambient-only white floor, black
    slab over x > 0.2, ortho camera on +z; optionally a 40%-opaque shell at z=1."""
    p = [(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, -2, 0), (2, 2, 0), (-2, 2, 0)]
    n = [(0, 0, 1)] * 6
    c = [(1, 1, 1, 1)] * 6
    z = 0.4
    p += [(0.2, -2, z), (4, -2, z), (4, 2, z), (0.2, -2, z), (4, 2, z), (0.2, 2, z)]
    n += [(0, 0, -1)] * 6
    c += [(0, 0, 0, 1)] * 6
    if shell:
        p += [(-2, -2, 1), (2, -2, 1), (2, 2, 1), (-2, -2, 1), (2, 2, 1), (-2, 2, 1)]
        n += [(0, 0, 1)] * 6
        c += [(0.5, 0.7, 0.9, 0.4)] * 6
    sc = Scene()
    sc.set_mesh(p, n, c, material=Material(ambient=1.0, diffuse=0.0))
    sc.camera = ortho_cam()
    sc.background = (0.1, 0.1, 0.1)
    return sc


@case
def coarse_ao_inline_neutrality():
    # test_ao_coarse.cpp:210-230: aoResDiv 0 and 1 are both the inline path
    sc = _coarse_scene()
    o0 = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0)
    o1 = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0, ao_res_div=1)
    e0 = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0, ao_multi_scale=True, ao_multibounce=True)
    e1 = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0, ao_multi_scale=True, ao_multibounce=True,
              ao_res_div=1)

    def chk(r):
        assert np.array_equal(r["o0"].color, r["o1"].color)
        assert np.array_equal(r["e0"].color, r["e1"].color)
    return Case("coarse_ao_inline_neutrality", {"o0": (sc, o0), "o1": (sc, o1), "e0": (sc, e0), "e1": (sc, e1)},
                chk)


@case
def coarse_ao_close_to_inline():
    # test_ao_coarse.cpp:232-269: "out" (= ss) coarse AO vs inline
    sc = _coarse_scene()
    full = opts(16, 16, supersample=3, ao_samples=256, ao_distance=100.0, ao_low_discrepancy=True)
    coarse = opts(16, 16, supersample=3, ao_samples=256, ao_distance=100.0, ao_low_discrepancy=True, ao_res_div=-1)
    f32 = opts(16, 16, supersample=3, ao_samples=32, ao_distance=100.0)
    c32 = opts(16, 16, supersample=3, ao_samples=32, ao_distance=100.0, ao_res_div=-1)

    def chk(r):
        assert np.abs(r["full"].color - r["coarse"].color).max() <= 0.05
        assert np.abs(r["f32"].color - r["c32"].color).mean() < 0.02
        assert not np.array_equal(r["full"].color, r["coarse"].color)
    return Case("coarse_ao_close_to_inline", {"full": (sc, full), "coarse": (sc, coarse), "f32": (sc, f32),
                                              "c32": (sc, c32)}, chk)


@case
def coarse_ao_fallback_exact():
    # test_ao_coarse.cpp:271-302: ss 1 with an explicit div 2: rejected
    # lookups fall back to the exact inline gather (same seeds), so those
    # pixels equal the inline render bitwise while interpolated ones do not
    sc = _coarse_scene()
    full = opts(24, 24, ao_samples=16, ao_distance=100.0)
    coarse = opts(24, 24, ao_samples=16, ao_distance=100.0, ao_res_div=2)

    def chk(r):
        same = np.all(r["full"].color == r["coarse"].color, axis=2)
        assert 0 < same.sum() < same.size, same.sum()
        assert r["coarse"].stats["ao_rays_traced"] + r["coarse"].stats["ao_rays_sampled"] < \
            r["full"].stats["ao_rays_traced"] + r["full"].stats["ao_rays_sampled"]
    return Case("coarse_ao_fallback_exact", {"full": (sc, full), "coarse": (sc, coarse)}, chk)


@case
def coarse_ao_transparency():
    # test_ao_coarse.cpp:346-373: a 40%-opaque shell over the floor
    sc = _coarse_scene(shell=True)
    full = opts(12, 12, supersample=2, ao_samples=16, ao_distance=100.0)
    coarse = opts(12, 12, supersample=2, ao_samples=16, ao_distance=100.0, ao_res_div=-1)

    def chk(r):
        assert np.abs(r["full"].color - r["coarse"].color).max() <= 0.05
    return Case("coarse_ao_transparency", {"full": (sc, full), "coarse": (sc, coarse)}, chk)


# ---------------------------------------------------------------- transparency
def blend_scene(quads, bg):
    """sceneOfBlend (test_render_transparency.cpp:53-66) with fragment alpha."""
    sc = Scene()
    P, N, Cc = [], [], []
    for color, z in quads:
        p, n, c = quad_mesh(color, z)
        P += p
        N += n
        Cc += c
    sc.set_mesh(P, N, Cc, material=Material.flat_outline())
    sc.camera = ortho_cam()
    sc.background = bg
    return sc


@case
def over_f1():
    # test_render_transparency.cpp:236-247
    sc = blend_scene([((1, 0, 0, 1.0), 0.0), ((0, 0, 1, 0.6), 1.0)], (0, 0, 0))

    def chk(r):
        approx(r["f"].color[C][0], 0.4, 1e-4)
        approx(r["f"].color[C][2], 0.6, 1e-4)
        approx(r["f"].color[C][3], 1.0, 1e-6)
    return Case("over_f1", {"f": (sc, opts())}, chk)


@case
def over_f2_order():
    # test_render_transparency.cpp:251-273
    a = blend_scene([((1, 0, 0, 1.0), 0.0), ((0, 0, 1, 0.5), 0.5), ((0, 1, 0, 0.5), 1.0)], (0, 0, 0))
    b = blend_scene([((1, 0, 0, 1.0), 0.0), ((0, 1, 0, 0.5), 0.5), ((0, 0, 1, 0.5), 1.0)], (0, 0, 0))

    def chk(r):
        approx(r["a"].color[C][1], 0.5, 1e-4)
        approx(r["a"].color[C][2], 0.25, 1e-4)
        approx(r["b"].color[C][2], 0.5, 1e-4)
        approx(r["b"].color[C][1], 0.25, 1e-4)
    return Case("over_f2_order", {"a": (a, opts()), "b": (b, opts())}, chk)


@case
def over_f3_nodedup():
    # test_render_transparency.cpp:277-288
    sc = blend_scene([((1, 0, 0, 1.0), 0.0), ((0, 0, 1, 0.5), 0.5), ((0, 0, 1, 0.5), 1.0)], (0, 0, 0))

    def chk(r):
        approx(r["f"].color[C][2], 0.75, 1e-4)
        approx(r["f"].color[C][0], 0.25, 1e-4)
    return Case("over_f3_nodedup", {"f": (sc, opts())}, chk)


@case
def over_f4_transparent_bg():
    # test_render_transparency.cpp:291-301
    sc = blend_scene([((0, 0, 1, 0.6), 1.0)], (0.2, 0.2, 0.2))

    def chk(r):
        approx(r["f"].color[C][2], 0.6, 1e-4)
        approx(r["f"].color[C][0], 0.0, 1e-4)
        approx(r["f"].color[C][3], 0.6, 1e-4)
    return Case("over_f4_transparent_bg", {"f": (sc, opts(transparent_background=True))}, chk)


# ---------------------------------------------------------------- curves
def cyl_scene(cyls, bg=(1, 1, 1)):
    sc = Scene()
    sc.camera = ortho_cam()
    sc.background = bg
    sc.ambient_color = (1, 1, 1)
    for c in cyls:
        sc.add_cylinder(**c)
    return sc


@case
def curve_c1_solid():
    # test_render_edges.cpp:34-56
    sc = cyl_scene([dict(p0=(-2, 0, 0), p1=(2, 0, 0), radius=0.5, color=(0, 0, 0, 1), open=True)])

    def chk(r):
        f = r["f"]
        assert f.color[C][0] < 0.05 and f.color[C][1] < 0.05 and f.color[C][2] < 0.05
        approx(f.color[C][3], 1.0, 1e-4)
    return Case("curve_c1_solid", {"f": (sc, opts())}, chk)


@case
def curve_c2_gradient():
    # test_render_edges.cpp:65-91
    sc = cyl_scene([dict(p0=(-2, 0, 0), p1=(2, 0, 0), radius=0.5, color=(0, 0, 0, 1), opacity1=0.0, open=True)])

    def chk(r):
        f = r["f"]
        lft, ctr, rgt = f.color[2, 1, 0], f.color[2, 2, 0], f.color[2, 3, 0]
        assert lft + 0.05 < ctr and ctr + 0.05 < rgt
        approx(ctr, 0.25, 0.02)
    return Case("curve_c2_gradient", {"f": (sc, opts(transparency=True))}, chk)


@case
def curve_c3_uniform():
    # test_render_edges.cpp:98-121
    sc = cyl_scene([dict(p0=(-2, 0, 0), p1=(2, 0, 0), radius=0.5, color=(0, 0, 0, 0.5), opacity1=-1.0, open=True)])

    def chk(r):
        f = r["f"]
        lft, ctr, rgt = f.color[2, 1, 0], f.color[2, 2, 0], f.color[2, 3, 0]
        approx(ctr, 0.25, 0.02)
        approx(lft, ctr, 0.02)
        approx(rgt, ctr, 0.02)
    return Case("curve_c3_uniform", {"f": (sc, opts(transparency=True))}, chk)


@case
def curve_c4_seam():
    # test_render_edges.cpp:134-160
    sc = cyl_scene([dict(p0=(-2, 0, 0), p1=(0, 0, 0), radius=0.5, color=(0, 0, 0, 0.5), open=True),
                    dict(p0=(0, 0, 0), p1=(2, 0, 0), radius=0.5, color=(0, 0, 0, 0.5), open=True)])

    def chk(r):
        f = r["f"]
        lft, ctr, rgt = f.color[2, 1, 0], f.color[2, 2, 0], f.color[2, 3, 0]
        approx(ctr, 0.25, 0.05)
        assert ctr + 0.03 >= lft and ctr + 0.03 >= rgt
    return Case("curve_c4_seam", {"f": (sc, opts(transparency=True))}, chk)


@case
def curve_c5_caps():
    # test_render_edges.cpp:172-205
    capped = cyl_scene([dict(p0=(-2, 0, 0), p1=(0, 0, 0), radius=0.45, color=(0, 0, 0, 1), open=False)])
    rnd = cyl_scene([dict(p0=(-2, 0, 0), p1=(0, 0, 0), radius=0.45, color=(0, 0, 0, 1), open=True)])

    def chk(r):
        fc, fo = r["capped"], r["round"]
        body, past = (5, 4), (5, 6)
        assert fc.color[body][0] < 0.1 and fo.color[body][0] < 0.1
        assert fc.color[past][0] > 0.8
        assert fo.color[past][0] + 0.2 < fc.color[past][0]
    return Case("curve_c5_caps", {"capped": (capped, opts(11, 11)), "round": (rnd, opts(11, 11))}, chk)


@case
def curve_c6_far_skip():
    # test_render_edges.cpp:206-233
    sc = Scene()
    sc.camera = ortho_cam(200.0)
    sc.background = (0, 0, 0)
    sc.ambient_color = (1, 1, 1)
    p, n, c = quad_mesh((1, 1, 1, 0.5))
    sc.set_mesh(p, n, c, material=Material.flat_outline())
    sc.add_sphere((0, 0, -5.001), 5.0, (0, 0, 1, 1), Material.flat_outline())
    sc.add_sphere((0, 0, -1.0), 0.5, (1, 0, 0, 1), Material.flat_outline())

    def chk(r):
        f = r["f"]
        assert f.color[C][2] > f.color[C][0] + 0.3
    return Case("curve_c6_far_skip", {"f": (sc, opts(transparency=True))}, chk)


# ---------------------------------------------------------------- clip planes
@case
def clip_planes():
    # test_clip_planes.cpp:20-77 (colour / view-z parts; the clip-cut G-buffer
    # belongs to the stroke-edge pass, which is outside this path)
    mat = Material(ambient=0.2, diffuse=0.8)

    def sc(near=-math.inf, far=math.inf):
        s = material_sphere_scene((1, 0, 0, 1), mat, (0, 0, 1))
        s.clip_near, s.clip_far = near, far
        return s
    sc_near = sc(near=10.0)
    sc_near.fog = Fog(enabled=False)

    def chk(r):
        f = r["none"]
        assert f.color[C][0] > 0.3 and f.color[C][2] < 0.5
        assert abs(r["far8"].color[C][2] - 1.0) <= 1e-4 and r["far8"].color[C][0] < 1e-4
        assert r["near12"].color[C][0] < 1e-4
        assert 10.5 < r["near10"].view_z[C] < 11.5
    return Case("clip_planes", {"none": (sc(), opts()), "far8": (sc(far=8.0), opts()),
                                "near12": (sc(near=12.0), opts()), "near10": (sc_near, opts())}, chk)


# ---------------------------------------------------------------- indexed mesh
@case
def mesh_indexed():
    # test_mesh_indexed.cpp:61-93 (indexed render == de-indexed, bitwise)
    verts = [(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, 2, 0)]
    idx = [0, 1, 2, 0, 2, 3]

    def mk(indexed):
        sc = Scene()
        if indexed:
            sc.set_mesh(verts, [(0, 0, 1)] * 4, [(0.5, 0.6, 0.7, 1)] * 4, index=np.array(idx).reshape(-1, 3),
                        material=Material(ambient=0.2, diffuse=0.8))
        else:
            sc.set_mesh([verts[i] for i in idx], [(0, 0, 1)] * 6, [(0.5, 0.6, 0.7, 1)] * 6,
                        material=Material(ambient=0.2, diffuse=0.8))
        sc.camera = ortho_cam()
        sc.background = (0.1, 0.2, 0.3)
        sc.lights = [DistantLight(direction=(0, 0, -1), intensity=0.5)]
        return sc
    o = opts(24, 24, supersample=2)

    def chk(r):
        assert np.array_equal(r["idx"].color, r["soup"].color)
    return Case("mesh_indexed", {"idx": (mk(True), o), "soup": (mk(False), o)}, chk)


# ---------------------------------------------------------------- fog in the frame
@case
def fog_frame():
    # fog.cpp:17-36 applied by pipeline.cpp:126-129 on the hi-res frame:
    # quad at view-z 10 with fog start 5, end 15 -> f = 0.5 toward fog colour
    sc = quad_scene(bg=(0, 0, 0))
    sc.fog = Fog(enabled=True, color=(1.0, 1.0, 1.0), start=5.0, end=15.0)

    def chk(r):
        f = r["f"]
        approx(f.color[C][0], 0.5 * 1.0 + 0.5 * 0.30, 1e-4)
        approx(f.color[C][2], 0.5 * 1.0 + 0.5 * 0.42, 1e-4)
        approx(r["bg"].color[C][0], 0.0, 1e-6)
    sc_bg = Scene()
    sc_bg.camera = ortho_cam()
    sc_bg.fog = Fog(enabled=True, color=(1, 1, 1), start=5.0, end=15.0)
    return Case("fog_frame", {"f": (sc, opts()), "bg": (sc_bg, opts())}, chk)


# ---------------------------------------------------------------- adaptive antialiasing
def half_quad_scene(x0=-2.0, x1=0.0) -> Scene:
    """makeQuadX / makeSceneX (test_adaptive_aa.cpp:87-126): a quad over
    [x0, x1] x [-2, 2] at z = 0 facing +Z, ortho camera of height 4, head-on
    key light 0.5, background 0.1."""
    sc = Scene()
    p = [(x0, -2, 0), (x1, -2, 0), (x1, 2, 0), (x0, -2, 0), (x1, 2, 0), (x0, 2, 0)]
    sc.set_mesh(p, [(0, 0, 1)] * 6, [(0.5, 0.6, 0.7, 1.0)] * 6, material=Material(ambient=0.2, diffuse=0.8))
    sc.camera = ortho_cam()
    sc.lights = [key_light()]
    sc.background = (0.1, 0.1, 0.1)
    return sc


@case
def aa_uniform():
    # test_adaptive_aa.cpp:361-379: nothing flags on a uniform view; the
    # replicated blocks equal the grid box average within interpolation rounding
    sc = half_quad_scene(-2.0, 2.0)

    def chk(r):
        g, a = r["grid"], r["adap"]
        assert a.color.shape == (6, 6, 4)
        assert np.abs(g.color - a.color).max() <= 1e-6
        assert a.aa_flagged == 0
    return Case("aa_uniform", {"grid": (sc, opts(6, 6, supersample=2)),
                               "adap": (sc, opts(6, 6, supersample=2, aa_mode=1))}, chk, tol=1e-6)


@case
def aa_silhouette():
    # test_adaptive_aa.cpp:381-437: flagged pixels re-shade every subpixel
    # exactly like the grid (bitwise after the resolve), the rest stays within
    # interpolation rounding; the mask (Wf x Hf) covers the silhouette column 4
    # and leaves columns 0 and 8 quiet
    sc = half_quad_scene()

    def chk(r):
        g, a = r["grid"], r["adap"]
        assert np.abs(g.color - a.color).max() <= 1e-6
        m = a.aa_mask
        assert m is not None and m.shape == (9, 9)
        flagged = m > 0
        assert flagged.sum() > 0 and a.aa_flagged == int(flagged.sum())
        assert np.array_equal(g.color[flagged], a.color[flagged])
        assert np.all(m[:, 4] > 0) and np.all(m[:, 0] == 0) and np.all(m[:, 8] == 0)
    return Case("aa_silhouette", {"grid": (sc, opts(9, 9, supersample=3)),
                                  "adap": (sc, opts(9, 9, supersample=3, aa_mode=1, aa_debug=True))}, chk, tol=1e-6)


@case
def aa_ao_compensation():
    # test_adaptive_aa.cpp:432-468: the unflagged floor under a slab takes
    # ONE centre evaluation with ss^2 x the AO samples; it estimates the same
    # openness as the grid's box average (within 0.05)
    sc = Scene()
    p, n, c = floor_slab_mesh(with_slab=True)
    p[6:] = [(0.2, -2, 0.4), (4, -2, 0.4), (4, 2, 0.4), (0.2, -2, 0.4), (4, 2, 0.4), (0.2, 2, 0.4)]
    c = [(0.5, 0.6, 0.7, 1.0)] * 6 + [(0, 0, 0, 1)] * 6
    sc.set_mesh(p, n, c, material=Material(ambient=1.0, diffuse=0.0))
    sc.camera = ortho_cam()
    sc.background = (0.1, 0.1, 0.1)
    grid = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0)
    adap = opts(8, 8, supersample=2, ao_samples=16, ao_distance=100.0, aa_mode=1)

    def chk(r):
        approx(r["grid"].color[4, 6, 0], r["adap"].color[4, 6, 0], 0.05)
    return Case("aa_ao_compensation", {"grid": (sc, grid), "adap": (sc, adap)}, chk)


@case
def aa_depth_equals_ss():
    # test_adaptive_aa.cpp:559-575: aaDepth == ss selects k == 1, the plain
    # subpixel path -- bitwise-equal to aaDepth 0
    sc = half_quad_scene()

    def chk(r):
        assert np.array_equal(r["d0"].color, r["d3"].color)
        assert np.array_equal(r["d0"].depth, r["d3"].depth)
    return Case("aa_depth_equals_ss", {"d0": (sc, opts(9, 9, supersample=3, aa_mode=1)),
                                       "d3": (sc, opts(9, 9, supersample=3, aa_mode=1, aa_depth=3))}, chk,
                tol=1e-6)


@case
def aa_depth_at_ss1():
    # test_adaptive_aa.cpp:577-608: ss 1 + aaDepth 2: unflagged pixels are the
    # plain one-sample render bitwise; flagged silhouette pixels are 2x2 box
    # averages that differ from the grid and lie between the two side colours
    sc = half_quad_scene()

    def chk(r):
        g, a = r["grid"], r["adap"]
        m = a.aa_mask
        assert m is not None and m.shape == (9, 9)
        flagged = m > 0
        same = np.all(g.color == a.color, axis=2)
        assert np.all(same[~flagged])
        diff = flagged & ~same
        assert diff.any()
        lo = np.minimum(g.color[4, 0, :3], g.color[4, 8, :3])
        hi = np.maximum(g.color[4, 0, :3], g.color[4, 8, :3])
        v = a.color[diff][:, :3]
        assert np.all(v >= lo - 1e-6) and np.all(v <= hi + 1e-6)
    return Case("aa_depth_at_ss1", {"grid": (sc, opts(9, 9)),
                                    "adap": (sc, opts(9, 9, aa_mode=1, aa_depth=2, aa_debug=True))}, chk, tol=1e-6)


# ---------------------------------------------------------------- pt1 GI
def _gi_opts(w, h, **kw):
    base = dict(gi=True, gi_integrator=1, pt1_half_res=False, pt1_denoise=False, shadows=True)
    base.update(kw)
    return RenderOptions(width=w, height=h, **base)


def gi_scene(ambient=(0.5, 0.5, 0.5)) -> Scene:
    """makeGiScene (test_pt1_render.cpp:27-66): a 0.7-grey floor quad and a
    real red sphere under one angled light, ambient 0.5."""
    sc = Scene()
    p = [(-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, -4), (4, 0, 4), (-4, 0, 4)]
    sc.set_mesh(p, [(0, 1, 0)] * 6, [(0.7, 0.7, 0.7, 1.0)] * 6, material=Material(ambient=0.2, diffuse=0.8))
    d = np.array([0.0, -0.35, -1.0])
    sc.camera = Camera(position=(0, 3, 8), direction=tuple(d / np.linalg.norm(d)), up=(0, 1, 0), fovy=40.0)
    ld = np.array([-0.4, -1.0, -0.3])
    sc.lights = [DistantLight(direction=tuple(ld / np.linalg.norm(ld)), color=(1, 1, 1), intensity=0.8)]
    sc.add_sphere((0, 1, 0), 1.0, (0.8, 0.3, 0.3, 1.0), Material(ambient=0.2, diffuse=0.8))
    sc.background = (0.1, 0.1, 0.15)
    sc.ambient_color = ambient
    return sc


@case
def gi_direct_parity():
    # test_pt1_render.cpp:143-208: at giIntensity 0 the composite adds +0, so
    # the frame is the direct shade with the receivers' constant ambient
    # dropped -- bitwise the plain render with ambient 0; at full strength
    # colour - indirect is that direct frame and every receiver gets indirect
    g0 = _gi_opts(48, 36, pt1_spp=4, gi_intensity=0.0)
    g1 = _gi_opts(48, 36, pt1_spp=4)
    plain = RenderOptions(width=48, height=36, shadows=True)

    def chk(r):
        assert np.array_equal(r["gi0"].color, r["amb0"].color)
        d = np.abs(r["gi1"].color[..., :3] - r["gi1"].indirect - r["gi0"].color[..., :3]).max()
        assert d <= 1e-5, d
        hit = r["gi1"].hit_id != 0xFFFFFFFF
        csg = hit & ((r["gi1"].hit_id >> 30) == 1)
        assert csg.sum() > 20 and np.all(r["gi1"].indirect[csg].max(axis=1) > 0.0)
        assert np.all(r["gi1"].indirect[~hit] == 0.0)
        assert r["gi1"].stats["gi_rays"] == 4 * int(hit.sum())  # one bounce: one query per path
    return Case("gi_direct_parity", {"gi0": (gi_scene(), g0), "amb0": (gi_scene((0.0, 0.0, 0.0)), plain),
                                     "gi1": (gi_scene(), g1)}, chk, tol=0.05)


def cornell_scene() -> Scene:
    """cornell::makeScene (test_pt1_render.cpp:85-135): open-top box, white
    floor and back wall, red left wall, green right wall, lit straight down."""
    white, red, green = (0.75, 0.75, 0.75, 1.0), (0.75, 0.05, 0.05, 1.0), (0.05, 0.75, 0.05, 1.0)
    quads = [(((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)), (0, 1, 0), white),
             (((-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)), (0, 0, 1), white),
             (((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)), (1, 0, 0), red),
             (((1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)), (-1, 0, 0), green)]
    P_, N_, C_ = [], [], []
    for (a, b, c_, d), n, col in quads:
        P_ += [a, b, c_, a, c_, d]
        N_ += [n] * 6
        C_ += [col] * 6
    sc = Scene()
    sc.set_mesh(P_, N_, C_, material=Material(ambient=0.2, diffuse=0.8))
    sc.camera = Camera(position=(0, 1, 3.2), direction=(0, 0, -1), up=(0, 1, 0), fovy=45.0)
    sc.lights = [DistantLight(direction=(0, -1, 0), color=(1, 1, 1), intensity=1.0)]
    sc.background = (0.0, 0.0, 0.0)
    sc.ambient_color = (0.25, 0.25, 0.25)
    return sc


@case
def gi_cornell_bleeding():
    # test_pt1_render.cpp:210-259: the floor next to the red wall picks up red
    # indirect (R > 1.15 G), mirrored for green; a second bounce raises the
    # floor-centre indirect
    sc = cornell_scene()

    def chk(r):
        ind, ind2 = r["b1"].indirect, r["b2"].indirect
        red, green, mid = ind[52, 6], ind[52, 57], ind[52, 32]
        assert red[0] > 1.15 * red[1] and red[0] > 0.0
        assert green[1] > 1.15 * green[0] and green[1] > 0.0
        assert np.all(mid > 0.0)
        assert np.all(ind2[52, 32] > mid)
    return Case("gi_cornell_bleeding", {"b1": (sc, _gi_opts(64, 64, pt1_spp=128)),
                                        "b2": (sc, _gi_opts(64, 64, pt1_spp=128, gi_bounces=2))}, chk, tol=0.05)


@case
def gi_isolated_sphere_no_self_occlusion():
    # test_pt1_render.cpp:261-310: the gather origin's epsilon scales with the
    # primary-ray distance, so a lone sphere seen from 200 units never
    # occludes its own hemisphere
    sc = Scene()
    sc.add_sphere((0, 0, 0), 1.0, (0.8, 0.3, 0.3, 1.0), Material(ambient=0.2, diffuse=0.8))
    sc.camera = Camera(position=(0, 0, 200), direction=(0, 0, -1), up=(0, 1, 0), orthographic=True, height=4.0)
    ld = np.array([-0.4, -0.3, -1.0])
    sc.lights = [DistantLight(direction=tuple(ld / np.linalg.norm(ld)), color=(1, 1, 1), intensity=0.8)]
    sc.background = (1.0, 1.0, 1.0)
    sc.ambient_color = (0.5, 0.5, 0.5)

    def chk(r):
        occ = r["f"].gi_occlusion
        assert occ[16, 16] < 0.05
        assert occ.max() < 0.2
    return Case("gi_isolated_sphere_no_self_occlusion", {"f": (sc, _gi_opts(33, 33, pt1_spp=16, shadows=False))},
                chk, tol=0.05)


@case
def gi_denoise():
    # pt1_denoise.cpp + denoise_atrous.cpp on the gathered E: the filtered
    # 2-spp indirect is much closer to a 256-spp gather than the raw one, and
    # 0 iterations leave E bitwise alone
    sc = gi_scene()

    def chk(r):
        ref = r["ref"].indirect
        recv = ref.max(axis=2) > 0
        err = lambda f: float(((f.indirect - ref)[recv] ** 2).mean())
        assert err(r["den"]) < 0.5 * err(r["raw"]), (err(r["den"]), err(r["raw"]))
        assert np.array_equal(r["it0"].color, r["raw"].color)
    o = dict(pt1_spp=2, pt1_ld=False)
    return Case("gi_denoise", {"raw": (sc, _gi_opts(96, 72, **o)),
                               "den": (sc, _gi_opts(96, 72, pt1_denoise=True, **o)),
                               "it0": (sc, _gi_opts(96, 72, pt1_denoise=True, denoise_iters=0, **o)),
                               "ref": (sc, _gi_opts(96, 72, pt1_spp=256))}, chk, tol=0.05)


@case
def gi_reduced_grid():
    # test_pt1_render.cpp:312-328 (gatherDiv 4: deterministic, indirect on the
    # floor, finite colour) and the reduced grid against the full-res gather:
    # at ss 2 the default divisor (= the supersample factor) gathers on a
    # half grid, upsamples, and re-gathers the rims; its indirect stays close
    # to the full-res gather's on average
    sc = gi_scene()
    d4 = _gi_opts(48, 36, pt1_spp=4, pt1_gather_div=4)
    half = _gi_opts(48, 36, supersample=2, pt1_spp=16, pt1_ld=True)
    full = _gi_opts(48, 36, supersample=2, pt1_spp=16, pt1_ld=True, pt1_gather_div=1)
    nopatch = _gi_opts(48, 36, supersample=2, pt1_spp=16, pt1_ld=True, pt1_edge_patch=False)

    def chk(r):
        a, b = r["d4a"], r["d4b"]
        assert np.array_equal(a.color, b.color) and np.array_equal(a.indirect, b.indirect)
        assert a.indirect.max() > 0.0 and np.isfinite(a.color).all()
        assert a.stats["gi_gbuffer_rays"] == 12 * 9
        recv = r["full"].indirect.max(axis=2) > 0
        d = np.abs(r["half"].indirect - r["full"].indirect)[recv].mean()
        assert d < 0.02, d
        assert r["half"].stats["gi_gbuffer_rays"] == 48 * 36
        # the rim patch re-gathers the unsupported pixels at full resolution
        assert r["half"].stats["gi_rays"] > r["nopatch"].stats["gi_rays"]
    return Case("gi_reduced_grid", {"d4a": (sc, d4), "d4b": (sc, d4), "half": (sc, half), "full": (sc, full),
                                    "nopatch": (sc, nopatch)}, chk, tol=0.05)


@case
def gi_pt2_default():
    # giIntegrator 2 with every reference default (blue-noise Sobol first
    # bounce, emission on, a-trous E denoise, gather grid = the supersample
    # factor): the same estimator as pt1 with another sampler, so its
    # indirect agrees with a converged pt1 gather on average
    sc = gi_scene()
    pt2 = RenderOptions(width=48, height=36, supersample=2, gi=True, pt1_spp=16, shadows=True)
    pt1 = _gi_opts(48, 36, supersample=2, pt1_spp=64, pt1_gather_div=1)

    def chk(r):
        recv = r["pt1"].indirect.max(axis=2) > 0
        d = np.abs(r["pt2"].indirect - r["pt1"].indirect)[recv].mean()
        assert d < 0.02, d
        assert np.array_equal(r["pt2"].color, r["pt2b"].color)
        assert r["pt2"].stats["gi_gbuffer_rays"] == 48 * 36
    return Case("gi_pt2_default", {"pt2": (sc, pt2), "pt2b": (sc, pt2), "pt1": (sc, pt1)}, chk, tol=0.05)
