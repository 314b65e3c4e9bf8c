"""Object-space silhouette edges (generateObjectSpaceEdges,
edges/object_space_edges.cpp) -- the reference's own checks
(test_object_space_edges.cpp:58-700) on the library's host code and on the
oracle, CPU only; the union clip on the device is in
tests/test_gpu_object_space_edges.py.

The library's sphere / cylinder silhouettes without the clip and its mesh
feature edges are host code: they run here with no device context. The
clip's point-in-solid tests run on the device, so the clip cases below hold
the oracle (its own brute-force restatement) to the reference's checks, and
the GPU tests hold the library to the oracle bit for bit.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import binding as ob
from umbreon_amd import capi
from umbreon_amd.edges import ObjectSpaceEdgeOptions, append_edges, generate_object_space_edges, object_space_edges
from umbreon_amd.scene import Camera, Scene


def ortho(direction=(0, 0, -1), position=(0, 0, 10)) -> Camera:
    return Camera(position=position, direction=direction, up=(0, 1, 0), orthographic=True, height=4.0)


def persp(position=(0, 0, 10)) -> Camera:
    return Camera(position=position, direction=(0, 0, -1), up=(0, 1, 0), fovy=40.0)


def opts(**kw) -> ObjectSpaceEdgeOptions:
    return ObjectSpaceEdgeOptions(enable=True, **kw)


def lib_edges(sc, o):
    """The library's edges (no device: only valid without a clip)."""
    return object_space_edges(sc, o, device=None)


def test_disabled_is_a_no_op():
    # (0)
    sc = Scene()
    sc.camera = ortho()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    out = generate_object_space_edges(sc, ObjectSpaceEdgeOptions())
    assert out is sc and len(out.cyl_p0) == 0


def _sphere_ortho_scene():
    sc = Scene()
    sc.camera = ortho()
    sc.add_sphere((1, 2, 3), 2.0, (1, 1, 1, 1), group=7)
    return sc


@pytest.mark.parametrize("who", ["library", "oracle"])
def test_sphere_ortho_great_circle_and_attributes(who):
    # (1): a ring of `segments` edges, every vertex r from the centre in the
    # plane across the view, a closed chain, open flat edges of the source group
    sc = _sphere_ortho_scene()
    o = opts(width=0.05, segments=32, color=(0.1, 0.2, 0.3), clip=False)
    e = lib_edges(sc, o) if who == "library" else ob.object_space_edges(sc, o)
    assert len(e["group"]) == 32
    O = np.array([1, 2, 3], np.float64)
    assert np.allclose(np.linalg.norm(e["p0"] - O, axis=1), 2.0, atol=1e-3)
    assert np.allclose(e["p0"][:, 2], 3.0, atol=1e-3)
    assert np.allclose(e["p1"], np.roll(e["p0"], -1, axis=0), atol=1e-6)  # closed chain
    assert (e["group"] == 7).all() and e["radius"] == pytest.approx(0.05)
    assert np.allclose(e["color"], [0.1, 0.2, 0.3, 1.0]) and (e["opacity1"] < 0).all()
    if who == "library":
        s2 = append_edges(sc, e)
        assert s2.cyl_open.all() and not s2.cyl_from_edge.any()
        m = s2.prim_materials[s2.cyl_material[0]]
        assert (m.ambient, m.diffuse, m.specular) == (1.0, 0.0, 0.0)


@pytest.mark.parametrize("who", ["library", "oracle"])
def test_sphere_perspective_horizon_circle(who):
    # (2): every vertex on the surface, its normal across the view ray
    sc = Scene()
    sc.camera = persp()
    sc.add_sphere((0, 0, 0), 2.0, (1, 1, 1, 1))
    o = opts(segments=24, clip=False)
    e = lib_edges(sc, o) if who == "library" else ob.object_space_edges(sc, o)
    assert len(e["group"]) == 24
    p = e["p0"].astype(np.float64)
    assert np.allclose(np.linalg.norm(p, axis=1), 2.0, atol=1e-3)
    n = p / np.linalg.norm(p, axis=1, keepdims=True)
    v = p - np.array([0, 0, 10.0])
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    assert np.abs((n * v).sum(axis=1)).max() <= 5e-3
    # (2b): the eye inside the sphere: no ring
    sc.camera = persp(position=(0, 0, 0))
    e = lib_edges(sc, o) if who == "library" else ob.object_space_edges(sc, o)
    assert len(e["group"]) == 0


@pytest.mark.parametrize("who", ["library", "oracle"])
def test_cylinder_side_lines_raise_and_end_on_rim(who):
    run = (lambda s, o: lib_edges(s, o)) if who == "library" else ob.object_space_edges
    # (3): two side lines at +-radius, across the axis and the view, parallel
    # to the axis, on opposite sides
    sc = Scene()
    sc.camera = ortho()
    sc.add_cylinder((-3, 0, 0), (3, 0, 0), 0.5, (1, 1, 1, 1), group=4)
    e = run(sc, opts(width=0.02, clip=False))
    assert len(e["group"]) == 2 and (e["group"] == 4).all()
    off = e["p0"] - np.stack([e["p0"][:, 0], np.zeros(2), np.zeros(2)], axis=1)
    assert np.allclose(np.linalg.norm(off, axis=1), 0.5, atol=1e-3)
    assert np.allclose(off[:, 0], 0.0) and np.allclose(off[0] + off[1], 0.0, atol=1e-3)
    d = e["p1"] - e["p0"]
    assert np.allclose(d[:, 1:], 0.0, atol=1e-3)
    # (3b): raise pushes the lines out by exactly `raise`
    e = run(sc, opts(raise_=0.1, clip=False))
    assert np.linalg.norm(e["p0"][0, 1:]) == pytest.approx(0.6, abs=1e-3)
    # (4): end-on, the rim of the nearer cap
    sc = Scene()
    sc.camera = ortho()
    sc.add_cylinder((0, 0, -2), (0, 0, 2), 0.7, (1, 1, 1, 1), group=9)
    e = run(sc, opts(clip=False))
    assert len(e["group"]) == 48 and (e["group"] == 9).all()
    assert np.allclose(np.linalg.norm(e["p0"][:, :2], axis=1), 0.7, atol=1e-3)
    assert np.allclose(e["p0"][:, 2], 2.0)  # the cap facing the eye


def test_snapshot_counts():
    # (5): 16 ring edges + 2 side lines, the appended edges not silhouetted
    sc = Scene()
    sc.camera = ortho()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    sc.add_cylinder((-2, 0, 0), (2, 0, 0), 0.3, (1, 1, 1, 1))
    out = generate_object_space_edges(sc, opts(segments=16, clip=False), device=None)
    assert len(out.cyl_p0) == 1 + 16 + 2


def test_library_matches_oracle_without_clip():
    # the host chords of both restatements agree bit for bit (perspective,
    # raise, end-on rims, a protein-like scene)
    from umbreon_amd.scenes import ball_and_stick_scene
    sc, _ = ball_and_stick_scene(n_atoms=300, width=64, height=36)
    sc.add_cylinder((0, 0, -1), (0, 0, 1), 0.3, (1, 1, 1, 1))
    for o in (opts(clip=False), opts(clip=False, raise_=0.05, segments=7)):
        a, b = lib_edges(sc, o), ob.object_space_edges(sc, o)
        for k in ("p0", "p1", "color", "opacity1", "group"):
            assert np.array_equal(a[k], b[k]), k


def _deep_inside(e, center, r):
    mid = 0.5 * (e["p0"] + e["p1"])
    return bool((np.linalg.norm(mid - np.asarray(center), axis=1) < r - 0.1).any())


def test_oracle_union_clip_trims_a_bond_through_an_atom():
    # (6)
    sc = Scene()
    sc.camera = ortho()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    sc.add_cylinder((-2, 0, 0), (2, 0, 0), 0.3, (1, 1, 1, 1))
    off = ob.object_space_edges(sc, opts(segments=16, clip=False))
    on = ob.object_space_edges(sc, opts(segments=16, clip=True))
    assert _deep_inside(off, (0, 0, 0), 1.0)
    assert not _deep_inside(on, (0, 0, 0), 1.0)
    assert len(on["group"]) > 0


def test_oracle_ball_and_stick_ring_rules():
    # (7): a thinner bond never clips the atom's ring, an equal one does
    def ring(cyl_r):
        sc = Scene()
        sc.camera = ortho()
        sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1), group=1)
        if cyl_r:
            sc.add_cylinder((0, 0, 0), (3, 0, 0), cyl_r, (1, 1, 1, 1), group=2)
        e = ob.object_space_edges(sc, opts(segments=24, clip=True))
        return int((e["group"] == 1).sum())
    full = ring(None)
    assert full == 24
    assert ring(0.3) == full
    assert ring(1.0) < full


# ---- mesh feature edges (library host code) ---------------------------------

def mesh_scene(pos, nrm=None) -> Scene:
    sc = Scene()
    sc.camera = ortho()
    pos = np.asarray(pos, np.float32)
    n = np.zeros_like(pos) if nrm is None else np.asarray(nrm, np.float32)
    sc.set_mesh(pos, n, np.ones((len(pos), 4), np.float32))
    return sc


def mesh_count(pos, nrm=None, **kw) -> int:
    base = dict(mesh_silhouette=False, mesh_crease=False, mesh_border=False)
    base.update(kw)
    return len(lib_edges(mesh_scene(pos, nrm), opts(**base))["group"])


TRI = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]


def test_mesh_border_silhouette_crease():
    # (8)
    assert mesh_count(TRI, [(0, 0, 1)] * 3, mesh_border=True) == 3
    assert mesh_count(TRI, [(0, 0, 1), (0, 0, 1), (0, 0, -1)], mesh_silhouette=True) == 1
    fold = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0), (0, 0, 1)]
    assert mesh_count(fold, [(0, 0, 1)] * 6, mesh_crease=True, crease_angle_deg=45.0) == 1
    assert mesh_count(fold, [(0, 0, 1)] * 6, mesh_crease=True, crease_angle_deg=120.0) == 0


def test_mesh_gates():
    # (9): convex-only, smooth-facet veto, border continuation veto
    convex = [(0, 0, 0), (2, 0, 0), (1, 1, -0.5), (2, 0, 0), (0, 0, 0), (1, -1, -0.5)]
    concave = [(0, 0, 0), (2, 0, 0), (1, 1, 0.5), (2, 0, 0), (0, 0, 0), (1, -1, 0.5)]
    assert mesh_count(convex, mesh_crease=True, crease_angle_deg=30.0) == 1
    assert mesh_count(convex, mesh_crease=True, crease_angle_deg=30.0, mesh_crease_convex_only=True) == 1
    assert mesh_count(concave, mesh_crease=True, crease_angle_deg=30.0, mesh_crease_convex_only=True) == 0
    facet = [(0, 0, 0), (1, 0, 0), (0.5, 1, 0.176), (1, 0, 0), (0, 0, 0), (0.5, -1, 0.176)]
    up = [(0, 0, 1)] * 6
    assert mesh_count(facet, up, mesh_crease=True, crease_angle_deg=10.0) == 1
    assert mesh_count(facet, up, mesh_crease=True, crease_angle_deg=10.0, mesh_crease_smooth_veto_deg=25.0) == 0
    fan = [(0, 0, 0), (1, 0, 0), (1.5, -1, 0), (1, 0, 0), (2, 0, 0), (1.5, -1, 0), (2, 0, 0), (3, 0, 0),
           (1.5, -1, 0)]
    b0 = mesh_count(fan, mesh_border=True)
    assert b0 == 5
    assert mesh_count(fan, mesh_border=True, mesh_border_coplanar_veto_deg=35.0) == b0 - 1


def test_mesh_on_contour_vertex_and_crease_degree():
    # (10)
    assert mesh_count(TRI, [(1, 0, 0), (0, 0, 1), (0, 0, -1)], mesh_silhouette=True) == 1
    cone = []
    ring = [(math.cos(i * math.pi / 3), math.sin(i * math.pi / 3), 0.0) for i in range(6)]
    for i in range(6):
        cone += [(0, 0, 1), ring[i], ring[(i + 1) % 6]]
    assert mesh_count(cone, mesh_crease=True, crease_angle_deg=20.0) == 6
    assert mesh_count(cone, mesh_crease=True, crease_angle_deg=20.0, mesh_crease_max_degree=4) == 0


def test_mesh_silhouette_lift_closed_form():
    # the smooth contour of a face whose vertex normals tilt from +z to -z:
    # the crossing at the n.v sign change, lifted by raise + width along the
    # interpolated normal and toward the eye by max(width / 2, 0.15 mean edge)
    w, raise_ = 0.04, 0.01
    nrm = [(0, 0, 1), (0, 0, 1), (0, 0, -1)]
    e = lib_edges(mesh_scene(TRI, nrm), opts(mesh_silhouette=True, mesh_crease=False, mesh_border=False, width=w,
                                             raise_=raise_))
    mean_edge = (1 + math.sqrt(2) + 1) / 3
    bias = max(0.5 * w, 0.15 * mean_edge)
    # crossings at the midpoints of edges (1,2) and (2,0); normals there are
    # 0 -> the normalized midpoint normal is the zero vector's fallback (itself)
    got = np.sort(np.concatenate([e["p0"], e["p1"]]), axis=0)
    want = np.sort(np.array([[0.5, 0.5, bias], [0.0, 0.5, bias]]), axis=0)
    assert np.allclose(got, want, atol=1e-6)


def test_visibility_clip_is_refused():
    with pytest.raises(capi.UmbreonError):
        lib_edges(mesh_scene(TRI, [(0, 0, 1)] * 3), opts(visibility_clip=True))


def test_clip_needs_a_device():
    sc = _sphere_ortho_scene()
    with pytest.raises(capi.UmbreonError):
        lib_edges(sc, opts(clip=True))
