"""ubg_pathtrace on the GPU against its oracle (oracle/oracle.cpp, the FP32
restatement with the same counter-based random numbers) and the closed forms
of tests/test_pt_oracle.py:

* parity on the path-traced molecular surface (1,500 atoms, GGX conductor and
  dielectric materials of roughness 0.05-0.6, 64 spherical lights, ReSTIR
  with 8 candidates and 4 neighbours, 4 bounces) at 160x90, 2 spp: per-pixel
  mean |d| <= 1e-3 and rays traced within 0.5 % (a path whose float rounding
  flips a Fresnel or a resampling choice follows another branch);
* the denoiser features equal their restatement from the oracle's first-hit
  record (the guides followed through dielectrics: at 99 % of pixels within
  1e-4), the guides through a glass sphere equal their closed form, and the
  denoised colour is the network's prediction on those features decoded
  (expm1(4 u) / 32); the noise descriptors equal 1 / sqrt(spp) and numpy's
  standard error over the per-sample values;
* white furnace of a smooth dielectric sphere (exactly 1), the Fresnel plane
  head-on, ReSTIR's mean against the all-lights estimator at 256 spp;
* multiple importance sampling of the lights: a glossy plane's direct light
  equals the quadrature over the light's cone;
* argument refusals.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import binding as ob
from umbreon_amd import capi
from umbreon_amd.pathtrace import CONDUCTOR, DIELECTRIC, DIFFUSE, PtMaterial, PtOptions, pathtrace
from umbreon_amd.scene import Camera, Scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def _built(device, sc):
    from umbreon_amd.render import BuiltScene
    return BuiltScene(device, sc, capi.BVH8)


@pytest.fixture(scope="module")
def surface():
    from umbreon_amd.scenes import pathtraced_surface_scene
    return pathtraced_surface_scene(n_atoms=1500, width=160, height=90, spp=2)


def test_parity_with_oracle(device, surface):
    sc, opt = surface
    bs = _built(device, sc)
    try:
        g = pathtrace(bs, sc, opt, want_features=True)
    finally:
        bs.close()
    c, first, rays = ob.pathtrace(sc, opt, precision=32, threads=16)
    d = np.abs(g.color[..., :3] - c[..., :3])
    hit = first[..., 7] > 0
    print(f"pathtrace parity: mean|d| {d.mean():.3g}, pixels |d| > 1e-3 {(d.max(-1) > 1e-3).mean():.4f}, "
          f"max {d.max():.3g}, rays gpu {g.rays} oracle {rays}, trace {g.trace_seconds * 1e3:.2f} ms", flush=True)
    assert hit.mean() > 0.1
    assert d.mean() <= 1e-3
    assert (d.max(-1) > 1e-3).mean() <= 0.02
    assert abs(g.rays / rays - 1) <= 5e-3
    assert np.all(g.color[..., 3] == 1.0)
    # features: encoded colour, albedo and normal guides, t / max t
    f = g.features
    # the colour encoded as log(1 + 32 c) / 4
    np.testing.assert_allclose(f[..., :3], 0.25 * np.log1p(32 * np.maximum(g.color[..., :3], 0)), rtol=1e-6,
                               atol=1e-7)
    depth = first[..., 3] / first[..., 3].max()
    np.testing.assert_allclose(f[..., 9], depth, rtol=1e-5, atol=1e-6)
    # the guides: a dielectric first hit's chain re-traverses the scene, so a
    # rounding difference can end it on another primitive at a few pixels
    dg = np.abs(f[..., 3:9] - first[..., [0, 1, 2, 4, 5, 6]]).max(-1)
    print(f"guides: mean|d| {dg.mean():.3g}, pixels |d| > 1e-4 {(dg > 1e-4).mean():.4f}", flush=True)
    assert (dg > 1e-4).mean() <= 0.01 and dg.mean() <= 2e-3


def test_noise_descriptors_are_the_sample_spread(device, surface):
    """Features 10-11: 1 / sqrt(n) and the standard error of the mean of the
    samples' (r + g + b) / 3 (unbiased variance), times the encoding's slope
    8 / (1 + 32 m) at that mean, at most 4. The samples of pass k do not
    depend on how many follow, so sample k's value is k c_k - (k - 1) c_(k-1)
    from the colours at k and k - 1 spp; numpy's variance of those values is
    the check."""
    sc, opt = surface
    n = 4
    bs = _built(device, sc)
    try:
        cols = [pathtrace(bs, sc, replace(opt, spp=k)).color[..., :3].astype(np.float64) for k in range(1, n)]
        g = pathtrace(bs, sc, replace(opt, spp=n), want_features=True)
    finally:
        bs.close()
    cols.append(g.color[..., :3].astype(np.float64))
    samples = [cols[0]] + [(k + 1) * cols[k] - k * cols[k - 1] for k in range(1, n)]
    lum = np.stack([s_.mean(-1) for s_ in samples])  # (n, H, W)
    m = lum.mean(0)
    se = np.sqrt(lum.var(0, ddof=1) / n)
    want = np.minimum(4.0, 8.0 * se / (1.0 + 32.0 * np.maximum(m, 0.0)))
    f = g.features
    assert np.all(f[..., 10] == np.float32(1.0 / np.sqrt(n)))
    d = np.abs(f[..., 11] - want)
    print(f"noise descriptor: mean {want.mean():.3g}, mean|d| {d.mean():.3g}, max {d.max():.3g}", flush=True)
    assert want.mean() > 1e-3  # a noisy frame (the background pixels have no spread)
    assert d.mean() <= 1e-4 and (d > 2e-3).mean() <= 1e-3


def test_denoised_colour_is_the_network_on_the_features(device, surface):
    from umbreon_amd.denoise import Unet
    sc, opt = surface
    bs = _built(device, sc)
    net = Unet(device, levels=5, channels=(32, 32, 32, 32, 32), seed=3)
    try:
        plain = pathtrace(bs, sc, opt, want_features=True)
        den = pathtrace(bs, sc, opt, denoise_net=net)
        m = 16
        H, W = opt.height, opt.width
        Hp, Wp = -(-H // m) * m, -(-W // m) * m
        feat = np.pad(plain.features, ((0, Hp - H), (0, Wp - W), (0, 0)), mode="edge")
        pred = net.infer(feat[None])[0, :H, :W]
        want = np.maximum(np.expm1(4 * pred) / 32, 0.0)
        np.testing.assert_allclose(den.color[..., :3], want, rtol=1e-5, atol=1e-6)
        assert den.denoise_seconds > 0
        # below 32 spp the default ensemble is the single pass
        one = pathtrace(bs, sc, replace(opt, denoise_ensemble=1), denoise_net=net)
        assert np.array_equal(one.color, den.color)
        # the self-ensemble: the mean of the network's predictions on the
        # turned features (pt_train._dihedral, normals turned with the pixels),
        # each turned back
        from umbreon_amd.pt_train import _dihedral

        def unturn(a, k):
            if k & 2:
                a = a[:, ::-1]
            if k & 1:
                a = a[:, :, ::-1]
            if k & 4:
                a = a.transpose(0, 2, 1, 3)
            return a

        for passes in (4, 8):
            ens = pathtrace(bs, sc, replace(opt, denoise_ensemble=passes), denoise_net=net)
            u = np.mean([unturn(net.infer(np.ascontiguousarray(_dihedral(feat[None], k, normals=True))), k)
                         for k in range(passes)], axis=0)[0, :H, :W]
            want = np.maximum(np.expm1(4 * u) / 32, 0.0)
            np.testing.assert_allclose(ens.color[..., :3], want, rtol=2e-5, atol=2e-6)
            assert not np.allclose(ens.color, den.color)
    finally:
        net.close()
        bs.close()


def test_guides_through_a_glass_sphere(device):
    """The closed form of tests/test_pt_oracle.py on the GPU: along the axis
    of a smooth glass sphere (ior 1.5, colour Cs) over a diffuse plane
    (colour Cp) the albedo guide is 0.04 + 0.96 Cs^2 Cp and the normal guide
    the plane's, (0, 0, 1) in camera space."""
    sc = Scene()
    v = np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32)
    sc.set_mesh(v, np.tile([0, 0, 1.0], (4, 1)).astype(np.float32),
                np.tile([0.3, 0.6, 0.9, 1.0], (4, 1)).astype(np.float32), np.array([0, 1, 2, 0, 2, 3], np.uint32))
    sc.add_spheres([[0, 0, 2]], [1.0], [[0.8, 0.5, 0.25, 1]])
    sc.camera = Camera(position=(0, 0, 10), direction=(0, 0, -1), up=(0, 1, 0), orthographic=True, height=0.01)
    opt = PtOptions(width=5, height=5, spp=1, bounces=1, prim_materials=[PtMaterial(DIELECTRIC, 0.0, 1.5)],
                    mesh_materials=[PtMaterial(DIFFUSE)])
    bs = _built(device, sc)
    try:
        f = pathtrace(bs, sc, opt, want_features=True).features
    finally:
        bs.close()
    want = 0.04 + 0.96 * np.array([0.8, 0.5, 0.25]) ** 2 * np.array([0.3, 0.6, 0.9])
    assert np.abs(f[..., 3:6] - want).max() <= 2e-5, f[2, 2, 3:6]
    assert np.abs(f[..., 6:9] - [0, 0, 1]).max() <= 1e-6


@pytest.mark.parametrize("mat", [PtMaterial(CONDUCTOR, 0.12), PtMaterial(CONDUCTOR, 0.5)])
def test_mis_direct_light_matches_quadrature(device, mat):
    """The lights reached by next-event estimation and by BSDF rays (balance
    heuristic) on the GPU: direct light of one spherical light on a glossy
    conductor plane equals the quadrature of f |cos| Le over the light's cone
    (tests/test_pt_oracle.py's case), and the image equals the oracle's."""
    from test_pt_oracle import _cone_quadrature, _glossy_under_light
    sc, opt, wo, c = _glossy_under_light(mat, spp=1024)
    bs = _built(device, sc)
    try:
        g = pathtrace(bs, sc, opt)
    finally:
        bs.close()
    want = _cone_quadrature(mat, wo, c)
    got = g.color[..., :3].reshape(-1, 3).mean(0)
    assert np.all(np.abs(got / want - 1) <= 0.02), (got, want)
    col, _, _ = ob.pathtrace(sc, opt, precision=32, threads=4)
    assert np.abs(g.color[..., :3] - col[..., :3]).max() <= 1e-4 * max(1.0, float(col.max()))


def test_white_furnace_dielectric_sphere(device):
    sc = Scene()
    sc.add_spheres([[0, 0, 0]], [1.0], [[1, 1, 1, 1]])
    sc.camera = Camera(position=(0, 0, 5), direction=(0, 0, -1), up=(0, 1, 0), fovy=30.0)
    opt = PtOptions(width=64, height=64, spp=16, bounces=64, env=(1, 1, 1),
                    prim_materials=[PtMaterial(DIELECTRIC, 0.0, 1.5)])
    bs = _built(device, sc)
    try:
        g = pathtrace(bs, sc, opt)
    finally:
        bs.close()
    assert np.abs(g.color[..., :3] - 1.0).max() <= 2e-5


def test_fresnel_plane(device):
    sc = Scene()
    v = np.array([[-6, -6, 0], [6, -6, 0], [6, 6, 0], [-6, 6, 0]], np.float32)
    sc.set_mesh(v, np.tile([0, 0, 1.0], (4, 1)), np.ones((4, 4)), np.array([0, 1, 2, 0, 2, 3], np.uint32))
    sc.camera = Camera(position=(0, 0, 10), direction=(0, 0, -1), up=(0, 1, 0), orthographic=True, height=2.0)
    opt = PtOptions(width=64, height=64, spp=256, bounces=1, env=(1, 1, 1),
                    mesh_materials=[PtMaterial(DIELECTRIC, 0.0, 1.5)])
    bs = _built(device, sc)
    try:
        g = pathtrace(bs, sc, opt)
    finally:
        bs.close()
    F = 0.04
    want = F + (1 - F) / 2.25
    se = np.sqrt(F * (1 - F)) * (1 - 1 / 2.25) / np.sqrt(64 * 64 * 256)
    assert abs(g.color[..., 0].mean() - want) <= 5 * se


def test_restir_mean_matches_all_lights(device):
    from umbreon_amd.scenes import pathtraced_surface_scene
    sc, opt = pathtraced_surface_scene(n_atoms=1500, width=192, height=108, spp=256)
    opt.prim_materials = [PtMaterial(m.type, 0.8, m.ior) for m in opt.prim_materials]
    opt.mesh_materials = [PtMaterial(DIFFUSE, 0.8, 1.5)]
    opt.bounces = 1
    bs = _built(device, sc)
    try:
        means = {}
        for name, K, all_lights in (("restir", 4, False), ("all", 0, True)):
            opt.restir_spatial, opt.nee_all_lights = K, all_lights
            means[name] = float(pathtrace(bs, sc, opt).color[..., :3].mean())
    finally:
        bs.close()
    print("ReSTIR vs all lights:", means, flush=True)
    assert abs(means["restir"] / means["all"] - 1) <= 0.01


def test_refusals(device, surface):
    sc, opt = surface
    from umbreon_amd.render import BuiltScene
    bs = _built(device, sc)
    b2 = BuiltScene(device, sc, capi.BVH2)
    try:
        import dataclasses
        for bad in (dict(restir_candidates=13), dict(restir_spatial=17), dict(bounces=0), dict(spp=0),
                    dict(lights=np.zeros((4097, 7), np.float32)),
                    dict(prim_materials=[PtMaterial(7, 0.3, 1.5)]), dict(denoise_ensemble=3)):
            with pytest.raises(capi.UmbreonError):
                pathtrace(bs, sc, dataclasses.replace(opt, **bad))
        with pytest.raises(capi.UmbreonError):
            pathtrace(b2, sc, opt)
    finally:
        b2.close()
        bs.close()
