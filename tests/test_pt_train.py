"""Host logic of the path-traced denoiser's training (umbreon_amd/pt_train.py):
the crop augmentation turns the camera-space normals with the image, and the
training windows copy the full frame's pixel footprint."""
from __future__ import annotations

import math

import numpy as np
import pytest

from umbreon_amd.pt_train import FULL_FOVY, FULL_H, SPP_STEPS, _dihedral, refine_crossings, window_scene


def _sphere_normals(n=33):
    yy, xx = np.mgrid[0:n, 0:n]
    c = (n - 1) / 2
    x, y = (xx - c) / c, -(yy - c) / c  # columns run along +right, rows along -up
    z = np.sqrt(np.maximum(0.0, 1.0 - x * x - y * y))
    f = np.zeros((1, n, n, 12))
    f[0, ..., 6], f[0, ..., 7], f[0, ..., 8] = x, y, z
    f[0, ..., 0] = xx  # a channel that is not a normal moves without changing sign
    return f


@pytest.mark.parametrize("k", range(8))
def test_dihedral_turns_normals_with_the_image(k):
    # the image of a centred sphere is the same under every flip / quarter
    # turn only if the normals turn with the pixels
    f = _sphere_normals()
    g = _dihedral(f, k, normals=True)
    assert np.allclose(g[..., 6:9], f[..., 6:9])
    assert g[..., 0].min() == 0 and sorted(np.unique(g[..., 0])) == sorted(np.unique(f[..., 0]))
    # without the normal handling a reflection breaks the sphere (k != 0)
    if k:
        assert not np.allclose(_dihedral(f, k)[..., 6:8], f[..., 6:8])


def test_window_scene_copies_the_full_frame_footprint():
    sc, opt = window_scene(1234, size=512)
    assert (opt.width, opt.height, opt.bounces, len(opt.lights)) == (512, 512, 4, 64)
    want = 2 * math.degrees(math.atan(math.tan(math.radians(FULL_FOVY / 2)) * 512 / FULL_H))
    assert abs(sc.camera.fovy - want) < 1e-9
    v = np.asarray(sc.positions, np.float64)
    c = 0.5 * (v.min(0) + v.max(0))
    r = np.linalg.norm(v - c, axis=1).max()
    dist = np.linalg.norm(np.asarray(sc.camera.position) - c)
    assert abs(dist - r / math.sin(math.radians(FULL_FOVY / 2)) * 1.02) < 1e-6 * dist
    d, up = np.asarray(sc.camera.direction), np.asarray(sc.camera.up)
    assert abs(np.linalg.norm(d) - 1) < 1e-9 and abs(np.linalg.norm(up) - 1) < 1e-9
    # aimed at a surface vertex: that vertex lies on the axis
    p = np.asarray(sc.camera.position)
    off = np.cross(v - p, d)
    assert np.linalg.norm(off, axis=1).min() < 1e-6 * dist
    # reproducible
    sc2, _ = window_scene(1234, size=512)
    assert sc2.camera == sc.camera and np.array_equal(sc2.positions, sc.positions)


def test_refine_crossings_finds_the_first_sample_count_at_each_threshold():
    # a PSNR that grows 3 dB per doubling from 24 dB at 1 spp: 30 dB at 4 spp
    # exactly (a step), 35 dB at 12.7 spp, 40 dB at 40.3 spp
    def measure(n):
        return (10.0 * n, 24.0 + 3.0 * math.log2(n))
    pts = {}
    for n in SPP_STEPS:
        pts[n] = measure(n)
        if pts[n][1] >= 40.0:
            break
    seen = []
    refine_crossings(pts, lambda n: (seen.append(n), measure(n))[1], log=lambda *_: None)
    # 30 dB: 3 -> 4, candidates 3.22, 3.46, 3.72 round to the steps: none
    # 35 dB: 12 -> 16, candidates 13, 14, 15: 13 reaches it (35.10 dB)
    # 40 dB: 32 -> 48, candidates 35, 39, 43 (32 * 1.5^(i/4)): 43 is the
    # first that reaches it (40.28 dB)
    assert seen == [13, 35, 39, 43]
    first = {db: min(n for n, p in pts.items() if p[1] >= db) for db in (30.0, 35.0, 40.0)}
    assert first == {30.0: 4, 35.0: 13, 40.0: 43}
    # a sweep whose first step already reaches a threshold is left alone
    pts = {1: (1.0, 41.0)}
    refine_crossings(pts, lambda n: (_ for _ in ()).throw(AssertionError(n)), log=lambda *_: None)
    assert list(pts) == [1]
