"""Object-space edges with the union clip on the device
(generateObjectSpaceEdges, edges/object_space_edges.cpp:151-229): the
reference's clip checks (test_object_space_edges.cpp:299-420) on the library,
the library's edges equal to the oracle's bit for bit (the point-in-solid
tests are exact float comparisons, so the runs must agree), and the render
integration (test_object_space_edges.cpp:566-700)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as ob
from test_object_space_edges import _deep_inside, opts, ortho, persp
from umbreon_amd.edges import generate_object_space_edges, object_space_edges
from umbreon_amd.scene import Camera, RenderOptions, Scene
from umbreon_amd.scenes import ball_and_stick_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def test_union_clip_trims_a_bond_through_an_atom(device):
    sc = Scene()
    sc.camera = ortho()
    sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1))
    sc.add_cylinder((-2, 0, 0), (2, 0, 0), 0.3, (1, 1, 1, 1))
    on = object_space_edges(sc, opts(segments=16, clip=True), device)
    assert not _deep_inside(on, (0, 0, 0), 1.0) and len(on["group"]) > 0


def test_ball_and_stick_ring_rules(device):
    def ring(cyl_r):
        sc = Scene()
        sc.camera = ortho()
        sc.add_sphere((0, 0, 0), 1.0, (1, 1, 1, 1), group=1)
        if cyl_r:
            sc.add_cylinder((0, 0, 0), (3, 0, 0), cyl_r, (1, 1, 1, 1), group=2)
        e = object_space_edges(sc, opts(segments=24, clip=True), device)
        return int((e["group"] == 1).sum())
    assert ring(None) == 24 and ring(0.3) == 24 and ring(1.0) < 24


@pytest.mark.parametrize("camera", ["perspective", "orthographic_raised"])
def test_clip_parity_ball_and_stick(device, camera):
    sc, _ = ball_and_stick_scene(n_atoms=600, width=64, height=36)
    o = opts(clip=True)
    if camera == "orthographic_raised":
        c = sc.camera
        sc.camera = Camera(position=c.position, direction=c.direction, up=c.up, orthographic=True, height=30.0)
        o = opts(clip=True, raise_=0.02, width=0.05, segments=20)
    print(f"object-space edges {camera}: {len(sc.sphere_center)} spheres, {len(sc.cyl_p0)} cylinders", flush=True)
    g = object_space_edges(sc, o, device)
    c = ob.object_space_edges(sc, o)
    print(f"edges GPU {len(g['group'])} oracle {len(c['group'])}", flush=True)
    assert len(g["group"]) > 1000
    for k in ("p0", "p1", "color", "opacity1", "group"):
        assert np.array_equal(g[k], c[k]), k


def test_render_integration(device):
    from umbreon_amd.render import BuiltScene

    def make():
        sc = Scene()
        sc.camera = Camera(position=(0, 0, 10), direction=(0, 0, -1), up=(0, 1, 0), orthographic=True, height=4.0)
        sc.background = (1, 1, 1)
        sc.add_sphere((0, 0, 0), 1.0, (0.3, 0.4, 0.5, 1.0))
        return sc

    o = opts(width=0.2, raise_=0.15)
    opt = RenderOptions(width=32, height=32)
    base = make()
    edged = generate_object_space_edges(base, o, device)
    assert len(base.cyl_p0) == 0 and len(edged.cyl_p0) == 48  # the caller's scene is untouched

    def gpu(sc):
        bs = BuiltScene(device, sc)
        try:
            return bs.render(sc, opt)
        finally:
            bs.close()

    fe, fp = gpu(edged), gpu(base)
    assert not np.array_equal(fe.color, fp.color)  # the black ring shows
    ce = ob.render(edged, opt)
    assert np.abs(fe.color - ce.color).max() <= 1e-5
