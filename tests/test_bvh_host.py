"""The library's BVH builders and traversal code, run on the CPU through the
host test hook (ubgx_host_query: same __host__ __device__ functions the gfx950
kernels execute), checked against the CPU oracle's brute force.

Both layouts -- full-precision BVH2 and the compressed 8-wide BVH with 8-bit
child boxes -- must return exactly the oracle's FP32 hits (same t, same
primitive) except at exact t ties, and identical occlusion answers.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as ob
from umbreon_amd import capi
from umbreon_amd.render import host_query
from umbreon_amd.scene import Scene
from test_oracle import _rays


def _mixed_scene(rng, n=150, open_chains=True):
    sc = Scene()
    c = rng.uniform(-4, 4, size=(n, 3))
    sc.add_spheres(c, rng.uniform(0.05, 0.6, n), np.ones((n, 4)))
    p0 = rng.uniform(-4, 4, size=(n, 3))
    p1 = p0 + rng.normal(size=(n, 3))
    sc.add_cylinders(p0, p1, rng.uniform(0.05, 0.3, n), np.ones((n, 4)), open=False)
    if open_chains:
        # polylines of open cylinders (chained round curves with joints)
        for k in range(12):
            pts = np.cumsum(rng.normal(scale=0.7, size=(6, 3)), axis=0) + rng.uniform(-3, 3, 3)
            sc.add_cylinders(pts[:-1], pts[1:], np.full(5, 0.12), np.ones((5, 4)), open=True)
    tri = rng.uniform(-4, 4, size=(3 * n, 3))
    tri[1::3] = tri[0::3] + rng.normal(scale=0.8, size=(n, 3))
    tri[2::3] = tri[0::3] + rng.normal(scale=0.8, size=(n, 3))
    sc.set_mesh(tri, np.tile([0, 0, 1], (3 * n, 1)), np.ones((3 * n, 4)))
    return sc


def _random_rays(rng, n, short=False):
    org = rng.uniform(-1.5, 1.5, size=(n, 3)) + np.array([0, 0, 14.0])
    d = rng.uniform(-4, 4, size=(n, 3)) - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if short:
        # AO-like: start inside the scene, short extent
        org = rng.uniform(-4, 4, size=(n, 3))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return _rays(org, d, tnear=1e-4, tfar=1.5)
    return _rays(org, d)


@pytest.mark.parametrize("layout", [capi.BVH2, capi.BVH8], ids=["bvh2", "bvh8"])
@pytest.mark.parametrize("short", [False, True], ids=["long", "short"])
def test_host_traversal_matches_oracle(layout, short):
    rng = np.random.default_rng(3 + int(short))
    sc = _mixed_scene(rng)
    r = _random_rays(rng, 6000, short)
    h, occ, st = host_query(sc, r, layout)
    ref = ob.intersect(sc, r, brute=True)
    same = (h["kind"] == ref["kind"]) & (h["prim_id"] == ref["prim_id"])
    assert same.mean() >= 0.9995, same.mean()
    hit = np.isfinite(ref["t"])
    assert np.array_equal(np.isfinite(h["t"]), hit)
    # identical fp32 formulas: t bit-identical (ties may pick another primitive at the same t)
    assert np.array_equal(h["t"][hit], ref["t"][hit])
    assert occ.max(initial=0) <= 1  # 2 = one-shot and resumable BVH8 any-hit traversals disagree
    assert np.array_equal(occ.astype(bool), hit)
    assert st["stack_depth"] >= 1 and st["node_visits"] > 0


def test_bvh8_is_shallower_and_visits_fewer_nodes():
    rng = np.random.default_rng(5)
    sc = _mixed_scene(rng, 2000, open_chains=False)
    r = _random_rays(rng, 4000)
    _, _, s2 = host_query(sc, r, capi.BVH2)
    _, _, s8 = host_query(sc, r, capi.BVH8)
    assert s8["node_bytes"] == 80 and s2["node_bytes"] == 64
    assert s8["depth"] < s2["depth"]
    assert s8["node_visits"] < 0.6 * s2["node_visits"]


def test_bvh8_large_scene_stack_bound():
    # 20k atom ball-and-stick: stack entries stay small and hits stay exact
    from umbreon_amd.scenes import ball_and_stick_scene
    sc, _ = ball_and_stick_scene(n_atoms=5000)
    rng = np.random.default_rng(9)
    c = sc.sphere_center[rng.integers(0, len(sc.sphere_center), 3000)]
    org = c + np.array([0, 0, 60.0])
    d = c - org + rng.normal(scale=0.3, size=c.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = _rays(org, d)
    h, occ, st = host_query(sc, r, capi.BVH8)
    ref = ob.intersect(sc, r)
    same = (h["kind"] == ref["kind"]) & (h["prim_id"] == ref["prim_id"])
    assert same.mean() >= 0.999
    assert st["stack_depth"] <= 16
