"""Batched rendering (ubg_render_batch): requests of different scenes, sizes
and options traced by the same launches give, request by request, the
images, AOVs and ray counts of single-request renders bit for bit -- the
determinism contract (output independent of parallel decomposition and
batch composition) -- whatever the tuning (AO waves per SIMD)."""
from __future__ import annotations

import numpy as np
import pytest

from umbreon_amd import capi
from umbreon_amd.scenes import ball_and_stick_scene, molecular_surface_scene, spheres_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def _requests():
    reqs = []
    sc, opt = spheres_scene(n_atoms=800, width=160, height=120)
    reqs.append((sc, opt))
    sc, opt = ball_and_stick_scene(n_atoms=3000, width=256, height=144, supersample=2)
    reqs.append((sc, opt))
    sc, opt = molecular_surface_scene(n_atoms=1500, width=200, height=150)
    reqs.append((sc, opt))
    sc, opt = ball_and_stick_scene(n_atoms=1200, width=96, height=96, seed=3)
    opt.ao_samples, opt.shadows = 0, True
    reqs.append((sc, opt))
    sc, opt = molecular_surface_scene(n_atoms=800, width=128, height=96, seed=5)
    opt.ao_samples = 4
    reqs.append((sc, opt))
    return reqs


def _same(a, b, tag):
    assert np.array_equal(a.color.view(np.uint32), b.color.view(np.uint32)), tag
    assert np.array_equal(a.depth.view(np.uint32), b.depth.view(np.uint32)), tag
    assert np.array_equal(a.view_z.view(np.uint32), b.view_z.view(np.uint32)), tag
    assert np.array_equal(a.hit_id, b.hit_id), tag
    assert np.array_equal(a.ao_open.view(np.uint32), b.ao_open.view(np.uint32)), tag
    for k in ("primary_rays", "shaded_hits", "ao_rays_traced", "ao_rays_sampled", "shadow_rays", "capped_rays"):
        assert a.stats[k] == b.stats[k], (tag, k)
    assert a.layers == b.layers, tag


@pytest.mark.parametrize("waves", [0, 2])
def test_batch_equals_single_renders(device, waves):
    from umbreon_amd.render import BuiltScene, render_batch
    reqs = _requests()
    built = [BuiltScene(device, sc) for sc, _ in reqs]
    try:
        singles = [b.render(sc, opt, want_ids=True, want_ao=True) for b, (sc, opt) in zip(built, reqs)]
        batch = render_batch(device, built, [r[0] for r in reqs], [r[1] for r in reqs], want_ids=True,
                             want_ao=True, ao_waves_per_simd=waves)
        for i, (a, b) in enumerate(zip(singles, batch)):
            print(f"request {i}: {a.width}x{a.height}, {a.layers} layers, {a.stats['primary_rays']} primary rays",
                  flush=True)
            _same(a, b, f"request {i}")
        # the same batch again (its layers replay the captured graphs, with
        # eager single renders in between) and without graphs
        for rep in range(2):
            again = render_batch(device, built, [r[0] for r in reqs], [r[1] for r in reqs], want_ids=True,
                                 want_ao=True, ao_waves_per_simd=waves)
            for i, (a, b) in enumerate(zip(singles, again)):
                _same(a, b, f"replay {rep} request {i}")
            b0 = built[rep].render(reqs[rep][0], reqs[rep][1], want_ids=True, want_ao=True)
            _same(singles[rep], b0, f"eager single render {rep}")
        direct = render_batch(device, built, [r[0] for r in reqs], [r[1] for r in reqs], want_ids=True,
                              want_ao=True, ao_waves_per_simd=waves, graph=False)
        for i, (a, b) in enumerate(zip(singles, direct)):
            _same(a, b, f"direct launches request {i}")
        # a different composition and order: the same images
        order = [4, 2, 0]
        sub = render_batch(device, [built[i] for i in order], [reqs[i][0] for i in order],
                           [reqs[i][1] for i in order], want_ids=True, want_ao=True)
        for j, i in enumerate(order):
            _same(singles[i], sub[j], f"reordered request {i}")
    finally:
        for b in built:
            b.close()


def test_leaf_size_does_not_change_images(device):
    """The BVH8 leaf size (an autotuner knob) changes the tree, not the
    images: first-hit ties resolve by (kind, primID) whatever the order."""
    from umbreon_amd.render import BuiltScene
    reqs = _requests()[1:3]
    for sc, opt in reqs:
        out = []
        for layout in (capi.BVH8, capi.BVH8_GPU_BUILD):
            for leaf in (1, 2, 3):
                b = BuiltScene(device, sc, layout, leaf_max=leaf)
                try:
                    out.append(b.render(sc, opt, want_ids=True, want_ao=True))
                finally:
                    b.close()
        for k, f in enumerate(out[1:]):
            _same(out[0], f, f"leaf/layout variant {k + 1}")


def test_batch_refuses_single_frame_options(device):
    from umbreon_amd.render import BuiltScene, render_batch
    sc, opt = spheres_scene(n_atoms=300, width=64, height=48)
    b = BuiltScene(device, sc)
    try:
        opt.gi, opt.gi_integrator = True, 2
        with pytest.raises(capi.UmbreonError) as e:
            render_batch(device, [b], [sc], [opt])
        assert e.value.status == capi.UBG_ERR_UNSUPPORTED
    finally:
        b.close()
