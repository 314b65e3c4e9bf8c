"""EmbreeRenderer::occluded (embree_renderer.cpp:294-357) batched: the edge
pass's segment visibility query with face exclusion, grazing and
coincident-plane rejection. Known answers on the oracle; the GPU path is held
to the oracle in tests/test_gpu_parity.py."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as ob
from umbreon_amd import capi
from umbreon_amd.scene import Material, Scene


def quad(instances=None):
    sc = Scene()
    p = [(-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, -2, 0), (2, 2, 0), (-2, 2, 0)]
    sc.set_mesh(p, [(0, 0, 1)] * 6, [(1, 1, 1, 1)] * 6, material=Material())
    if instances is not None:
        sc.instance_offsets = np.asarray(instances, np.float32)
    return sc


def seg(p, q, first=0, count=0):
    s = np.zeros(1, capi.segment_dtype())
    s["p"], s["q"], s["exclude_first"], s["exclude_count"] = p, q, first, count
    return s


def occ(sc, s, **kw):
    return int(ob.occluded_segments(sc, s, **kw)[0])


def test_plain_crossing_and_trim():
    sc = quad()
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, -1))) == 1
    # the segment's end on the surface is trimmed away (tfar = len * (1 - eps))
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, 0))) == 0
    # and so is its start (tnear = max(eps * len, eps))
    assert occ(sc, seg((0.5, 0.3, 0), (0.5, 0.3, -1))) == 0
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, 1))) == 0  # zero length


def test_face_exclusion():
    sc = quad()
    # (0.5, 0.3) lies in triangle 0 ((-2,-2), (2,-2), (2,2)): x > y
    s = seg((0.5, 0.3, 1), (0.5, 0.3, -1), 0, 1)
    assert occ(sc, s, exclude_faces=[0]) == 0
    assert occ(sc, s, exclude_faces=[1]) == 1
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, -1), 1, 2), exclude_faces=[7, 1, 0]) == 0
    # exclusion applies to mesh faces only: a sphere still occludes
    sc.add_sphere((0, 0, 5), 0.5, (1, 1, 1, 1))
    assert occ(sc, seg((0, 0, 7), (0, 0, 3), 0, 2), exclude_faces=[0, 1]) == 1


def test_instance_copies_map_to_base_faces():
    sc = quad(instances=[(0, 0, 0), (0, 0, -3)])
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, -5))) == 1
    # base face 0 excludes it in both copies
    assert occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, -5), 0, 1), exclude_faces=[0]) == 0
    # an exclusion range beyond the list is an error, not a silent clamp
    with pytest.raises(capi.UmbreonError):
        occ(sc, seg((0.5, 0.3, 1), (0.5, 0.3, -5), 0, 1))


def test_grazing_rejection():
    sc = quad()
    # crossing the plane at 1 degree through the origin: |cos(dir, Ng)| = sin(1 deg) ~ 0.01745
    a = np.radians(1.0)
    u = np.array([np.cos(a), 0.0, -np.sin(a)])
    s = seg(tuple(-u), tuple(u))
    assert occ(sc, s) == 1
    assert occ(sc, s, graze_cos_eps=0.02) == 0
    assert occ(sc, s, graze_cos_eps=0.015) == 1


def test_coplanar_rejection():
    sc = quad()
    s = seg((0.5, 0.3, 1e-3), (0.5, 0.3, -1))
    assert occ(sc, s) == 1
    # the hit plane passes 1e-3 from the origin: coincident for eps 1e-2
    assert occ(sc, s, coplanar_eps=1e-2) == 0
    assert occ(sc, s, coplanar_eps=1e-4) == 1


def test_sphere_next_root_after_rejection():
    sc = Scene()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    # from just outside the surface through the sphere: the near root is a
    # coincident hit (rejected), the far root 2 units on counts
    s = seg((0, 0, 1.0 + 1e-3), (0, 0, -3))
    assert occ(sc, s) == 1
    assert occ(sc, s, coplanar_eps=1e-2) == 1
    # stopping inside: only the rejected near root lies on the segment
    assert occ(sc, seg((0, 0, 1.0 + 1e-3), (0, 0, 0)), coplanar_eps=1e-2) == 0


def test_unfiltered_equals_occluded_query():
    from test_oracle import _random_prim_scene, _rays
    rng = np.random.default_rng(2)
    sc = _random_prim_scene(rng, 150)
    n = 3000
    p = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
    q = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
    s = np.zeros(n, capi.segment_dtype())
    s["p"], s["q"] = p, q
    o1 = ob.occluded_segments(sc, s)
    d = (q - p).astype(np.float32)
    L = np.sqrt((d * d).sum(axis=1, dtype=np.float32)).astype(np.float32)
    ok = L > 0
    dirs = (d / L[:, None]).astype(np.float32)
    r = _rays(p, dirs)
    r["tnear"] = np.maximum(np.float32(1e-4) * L, np.float32(1e-4))
    r["tfar"] = L * np.float32(1.0 - 1e-4)
    o2 = ob.occluded(sc, r, brute=True)
    assert np.array_equal(o1[ok], o2[ok])
    assert 0.05 < o1.mean() < 0.95
