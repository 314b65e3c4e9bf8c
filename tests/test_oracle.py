"""Pins for the CPU oracle (oracle/oracle.cpp) -- runs without a GPU.

What fixes each oracle function:
* integratePixel / shadeHit / shadeLocal / computeShadow / computeAO / fog /
  downsample / gamma: the reference's own closed-form render tests
  (tests/refcases.py, each citing its source assertion), in FP32 and FP64.
* TEA-8 (secondary_rays.hpp:51-58): the published TEA round function with the
  reference's keys, checked against an independent bit-level implementation
  written from Wheeler & Needham's description plus uniformity of the derived
  [0,1) samples (a wrong shift or key fails both).
* Intersectors: closed forms (sphere near/far roots, cylinder body and cap
  distances, triangle t = 1 / primID 0 of test_pt1_embree.cpp:64-77) and FP64
  brute force on random rays (FP32 hit t within 1e-5 relative except at
  grazing incidence).
* BVH: identical hits to brute force on random scenes.
* Curve chaining (curve_build.cpp:88-199): hand-derived chains for straight,
  branched, cyclic, group- and opacity-incompatible polylines.
* Post-process: test_fog.cpp:31-97 and test_image_ops.cpp:23-93 values.
* Ray accounting: counts a scene fixes exactly (e.g. 25 primary + 25*16 AO).
Parity unpinned (no live Embree here): exact curve t/u/Ng, triangle edge
watertightness, interpolation rounding and the order among equal-t hits.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import binding as ob
from refcases import CASES
from umbreon_amd.scene import Camera, DistantLight, Material, RenderOptions, Scene


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
@pytest.mark.parametrize("precision", [32, 64])
def test_reference_closed_forms(case, precision):
    res = {k: ob.render(s, o, precision=precision, threads=2) for k, (s, o) in case.renders.items()}
    case.check(res)


# ---------------------------------------------------------------- TEA
def _tea_indep(v0: int, v1: int, rounds: int = 8):
    # TEA encipher (Wheeler & Needham 1994) with key (A341316C, C8013EA4,
    # AD90777D, 7E95761E), delta 0x9E3779B9, written per the published round
    # function: v0 += ((v1<<4)+k0) ^ (v1+sum) ^ ((v1>>5)+k1), then v1 likewise.
    M = 0xFFFFFFFF
    k = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + k[0]) & M ^ ((v1 + s) & M) ^ (((v1 >> 5) + k[1]) & M))) & M
        v1 = (v1 + ((((v0 << 4) & M) + k[2]) & M ^ ((v0 + s) & M) ^ (((v0 >> 5) + k[3]) & M))) & M
    return v0, v1


def test_tea_matches_published_round_function():
    rng = np.random.default_rng(1)
    for a, b in rng.integers(0, 2**32, size=(200, 2), dtype=np.uint64):
        assert ob.tea2(int(a), int(b)) == _tea_indep(int(a), int(b))
    assert ob.tea2(0, 0) == _tea_indep(0, 0)


def test_tea_unorm_uniform():
    vals = []
    for px in range(64):
        for i in range(64):
            a, b = ob.tea2(px, i)
            vals += [(a >> 8) * 2.0**-24, (b >> 8) * 2.0**-24]
    v = np.array(vals)
    assert v.min() >= 0.0 and v.max() < 1.0
    hist, _ = np.histogram(v, bins=16, range=(0, 1))
    exp = len(v) / 16
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


# ---------------------------------------------------------------- intersectors
def _rays(orgs, dirs, tnear=0.0, tfar=np.inf):
    from umbreon_amd import capi
    r = np.zeros(len(orgs), dtype=capi.ray_dtype())
    r["org"] = orgs
    r["dir"] = dirs
    r["tnear"] = tnear
    r["tfar"] = tfar
    return r


def test_triangle_smoke():
    # test_pt1_embree.cpp:64-77: unit triangle, ray from (0.25,0.25,1) down
    sc = Scene()
    sc.set_mesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [(0, 0, 1)] * 3, [(1, 1, 1, 1)] * 3)
    h = ob.intersect(sc, _rays([(0.25, 0.25, 1.0), (2, 2, 1)], [(0, 0, -1), (0, 0, -1)]))
    assert h["kind"][0] == 0 and h["prim_id"][0] == 0
    assert abs(h["t"][0] - 1.0) < 1e-5
    assert abs(h["u"][0] - 0.25) < 1e-6 and abs(h["v"][0] - 0.25) < 1e-6
    assert h["ng"][0][2] > 0  # (v1-v0)x(v2-v0) = +z
    assert h["kind"][1] == 0xFFFFFFFF


def test_sphere_roots():
    sc = Scene()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    # from outside: near root; from inside (tnear past the front): far root
    r = _rays([(0, 0, 10), (0, 0, 10), (0, 0, 0)], [(0, 0, -1)] * 3, tnear=np.array([0.0, 9.5, 0.0]))
    for prec in (32, 64):
        h = ob.intersect(sc, r, precision=prec)
        assert abs(h["t"][0] - 9.0) < 1e-5
        assert abs(h["t"][1] - 11.0) < 1e-5
        assert abs(h["t"][2] - 1.0) < 1e-6
        assert h["ng"][0][2] > 0 and h["ng"][1][2] < 0


def test_cone_body_and_caps():
    # capped cylinder along x from -1 to 1, radius 0.5
    sc = Scene()
    sc.add_cylinder((-1, 0, 0), (1, 0, 0), 0.5, (1, 1, 1, 1), open=False)
    r = _rays([(0.3, 0, 5), (-5, 0, 0.1), (5, 0.2, 0), (1.2, 0, 5)],
              [(0, 0, -1), (1, 0, 0), (-1, 0, 0), (0, 0, -1)])
    for prec in (32, 64):
        h = ob.intersect(sc, r, precision=prec)
        assert abs(h["t"][0] - 4.5) < 1e-5 and abs(h["u"][0] - 0.65) < 1e-6   # body
        assert abs(h["t"][1] - 4.0) < 1e-5 and h["u"][1] == 0.0 and h["ng"][1][0] < 0  # cap at p0
        assert abs(h["t"][2] - 4.0) < 1e-5 and h["u"][2] == 1.0 and h["ng"][2][0] > 0  # cap at p1
        assert h["kind"][3] == 0xFFFFFFFF  # past the flat cap: miss


def test_round_capsule_bulges_past_endpoint():
    sc = Scene()
    sc.add_cylinder((-1, 0, 0), (1, 0, 0), 0.5, (1, 1, 1, 1), open=True)
    r = _rays([(1.2, 0, 5)], [(0, 0, -1)])
    h = ob.intersect(sc, r)
    # sphere cap at p1: z = sqrt(0.25 - 0.04)
    assert h["kind"][0] == 2 and abs(h["t"][0] - (5 - math.sqrt(0.21))) < 1e-5


def _random_prim_scene(rng, n=60):
    sc = Scene()
    c = rng.uniform(-3, 3, size=(n, 3))
    sc.add_spheres(c, rng.uniform(0.1, 0.6, n), np.ones((n, 4)))
    p0 = rng.uniform(-3, 3, size=(n, 3))
    p1 = p0 + rng.normal(size=(n, 3))
    sc.add_cylinders(p0, p1, rng.uniform(0.05, 0.3, n), np.ones((n, 4)), open=False)
    tri = rng.uniform(-3, 3, size=(3 * n, 3))
    sc.set_mesh(tri, np.tile([0, 0, 1], (3 * n, 1)), np.ones((3 * n, 4)))
    return sc


def test_fp32_vs_fp64_and_bvh_vs_brute():
    rng = np.random.default_rng(7)
    sc = _random_prim_scene(rng)
    n = 4000
    org = rng.uniform(-1, 1, size=(n, 3)) + np.array([0, 0, 12.0])
    tgt = rng.uniform(-3, 3, size=(n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = _rays(org, d)
    hb = ob.intersect(sc, r, brute=True)
    hv = ob.intersect(sc, r)
    h64 = ob.intersect(sc, r, precision=64, brute=True)
    # BVH == brute force (same precision): identical t; ids identical except at exact ties
    assert np.array_equal(hb["t"], hv["t"])
    same = (hb["kind"] == hv["kind"]) & (hb["prim_id"] == hv["prim_id"])
    assert same.mean() > 0.999
    # FP32 vs FP64 brute force
    hit = np.isfinite(h64["t"]) & np.isfinite(hb["t"])
    agree_id = (hb["kind"] == h64["kind"]) & (hb["prim_id"] == h64["prim_id"])
    assert agree_id.mean() > 0.995
    rel = np.abs(hb["t"][hit & agree_id] - h64["t"][hit & agree_id]) / h64["t"][hit & agree_id]
    assert np.quantile(rel, 0.999) < 1e-5
    occ = ob.occluded(sc, r)
    assert np.array_equal(occ.astype(bool), np.isfinite(hb["t"]))


# ---------------------------------------------------------------- chaining
def _open_chain_scene(segs, groups=None, op1=None):
    sc = Scene()
    for i, (a, b) in enumerate(segs):
        sc.add_cylinder(a, b, 0.1, (0, 0, 0, 1), open=True, group=(groups[i] if groups else 0),
                        opacity1=(op1[i] if op1 else -1.0))
    return sc


def test_chain_straight_and_reversed_order():
    # three collinear segments given out of order: chain starts at the one with no predecessor
    segs = [((1, 0, 0), (2, 0, 0)), ((0, 0, 0), (1, 0, 0)), ((2, 0, 0), (3, 0, 0))]
    src, fl = ob.chain_curves(_open_chain_scene(segs))
    assert list(src) == [1, 0, 2]
    assert list(fl) == [2, 3, 1]  # first: RIGHT, middle: LEFT|RIGHT, last: LEFT


def test_chain_branch_breaks_at_degree3():
    segs = [((0, 0, 0), (1, 0, 0)), ((1, 0, 0), (2, 0, 0)), ((1, 0, 0), (1, 1, 0))]
    src, fl = ob.chain_curves(_open_chain_scene(segs))
    assert sorted(src) == [0, 1, 2] and list(fl) == [0, 0, 0]


def test_chain_cycle_and_incompatible():
    cyc = [((0, 0, 0), (1, 0, 0)), ((1, 0, 0), (1, 1, 0)), ((1, 1, 0), (0, 0, 0))]
    src, fl = ob.chain_curves(_open_chain_scene(cyc))
    assert list(src) == [0, 1, 2] and list(fl) == [2, 3, 1]
    segs = [((0, 0, 0), (1, 0, 0)), ((1, 0, 0), (2, 0, 0))]
    src, fl = ob.chain_curves(_open_chain_scene(segs, groups=[0, 1]))
    assert list(fl) == [0, 0]
    src, fl = ob.chain_curves(_open_chain_scene(segs, op1=[-1.0, 0.5]))
    assert list(fl) == [0, 0]
    # endpoints within the 1e-4 hash cell still chain
    segs = [((0, 0, 0), (1, 0, 0)), ((1.00000001, 0, 0), (2, 0, 0))]
    src, fl = ob.chain_curves(_open_chain_scene(segs))
    assert list(fl) == [2, 1]


# ---------------------------------------------------------------- post-process
def test_fog_factor():
    # test_fog.cpp:31-54
    assert ob.fog_factor(True, 100, 300, 100) == 1.0
    assert ob.fog_factor(True, 100, 300, 50) == 1.0
    assert ob.fog_factor(True, 100, 300, 300) == 0.0
    assert ob.fog_factor(True, 100, 300, 500) == 0.0
    assert abs(ob.fog_factor(True, 100, 300, 200) - 0.5) < 1e-6
    assert ob.fog_factor(False, 100, 300, 200) == 1.0
    assert ob.fog_factor(True, 300, 300, 299) == 1.0
    assert ob.fog_factor(True, 300, 300, 301) == 0.0


def test_box_downsample_and_srgb():
    # test_image_ops.cpp:23-46, 74-93
    src = np.arange(16, dtype=np.float32).reshape(4, 4, 1)
    out = ob.box_downsample(src, 2)
    assert np.allclose(out[:, :, 0], [[2.5, 4.5], [10.5, 12.5]])
    src = np.array([np.nan, 2.0, 4.0, np.inf], np.float32).reshape(2, 2, 1)
    assert abs(ob.box_downsample(src, 2)[0, 0, 0] - 1.5) < 1e-6
    assert int(ob.srgb_encode(0.5) * 255 + 0.5) == 188
    assert int(ob.srgb_encode(1.0) * 255 + 0.5) == 255
    assert int(ob.srgb_encode(2.0) * 255 + 0.5) == 255
    assert abs(ob.srgb_decode(ob.srgb_encode(0.3)) - 0.3) < 1e-5


# ---------------------------------------------------------------- ray accounting
def test_ray_counts_closed_form():
    from refcases import quad_scene
    sc = quad_scene(bg=(0, 0, 0))
    f = ob.render(sc, RenderOptions(width=5, height=5, ao_samples=16, ao_distance=100.0, shadows=True))
    s = f.stats
    assert s["primary_rays"] == 25 and s["shaded_hits"] == 25
    assert s["ao_rays_traced"] + s["ao_rays_sampled"] == 25 * 16
    assert s["shadow_rays"] == 25  # one light with N.L > 0
    # transparency walk: 3 layers (two 0.5 quads over opaque) -> 3 primary rays per pixel
    from refcases import blend_scene
    sc = blend_scene([((1, 0, 0, 1.0), 0.0), ((0, 0, 1, 0.5), 0.5), ((0, 1, 0, 0.5), 1.0)], (0, 0, 0))
    f = ob.render(sc, RenderOptions(width=5, height=5))
    assert f.stats["primary_rays"] == 75 and f.stats["shaded_hits"] == 75
    # layer cap: max 1 transparent layer -> ray capped after the first fragment
    f = ob.render(sc, RenderOptions(width=5, height=5, max_transparent_layers=1))
    assert f.stats["capped_rays"] == 25 and f.stats["primary_rays"] == 50


def test_fp32_matches_fp64_on_small_protein():
    from umbreon_amd.scenes import spheres_scene
    sc, opt = spheres_scene(n_atoms=300, width=96, height=72)
    a = ob.render(sc, opt, precision=32, threads=4)
    b = ob.render(sc, opt, precision=64, threads=4)
    same = a.hit_id == b.hit_id
    assert same.mean() > 0.999
    d = np.abs(a.color - b.color).max(axis=2)
    assert np.quantile(d[same], 0.99) < 1e-4
