"""The serving layer's host logic (CPU): the seeded request stream, the
per-request camera and the HBM-budget admission."""
from __future__ import annotations

import math

import numpy as np

from umbreon_amd import serve
from umbreon_amd.scenes import spheres_scene


def test_request_stream_shape():
    s = serve.request_stream(256, rate=40.0, seed=0)
    assert len(s) == 256 and [r.rid for r in s] == list(range(256))
    gaps = np.diff([0.0] + [r.arrival for r in s])
    assert (gaps > 0).all()
    assert abs(gaps.mean() - 1 / 40.0) < 0.25 / 40.0  # exponential, mean 25 ms
    pool = serve.pool_sizes()
    assert pool[0] == 500 and pool[-1] == 200_000
    assert {r.atoms for r in s} <= set(pool)
    assert all(r.atoms <= 40_000 for r in s if r.rep == "surface")
    assert {(r.width, r.height) for r in s} == set(serve.SIZES)
    assert {r.rep for r in s} == set(serve.REPS)
    # log-uniform sizes: every pool size used, small and large alike
    counts = np.array([sum(r.atoms == p for r in s if r.rep != "surface") for p in pool])
    assert (counts > 0).all()
    assert serve.request_stream(256, seed=0)[17] == s[17]  # seeded


def test_request_view_orbits_the_framing_camera():
    sc, _ = spheres_scene(n_atoms=300, width=64, height=48)
    r = serve.Request(0, 0.0, 300, "spheres", 512, 512, math.pi / 2)
    v, opt = serve.request_view(sc, r)
    p0, p1 = np.array(sc.camera.position), np.array(v.camera.position)
    c = p0 - np.array([0.0, 0.0, serve._framing_distance(sc)])
    assert abs(np.linalg.norm(p1 - c) - np.linalg.norm(p0 - c)) < 1e-4 * np.linalg.norm(p0 - c)
    d = np.array(v.camera.direction)
    assert np.allclose(d / np.linalg.norm(d), (c - p1) / np.linalg.norm(c - p1), atol=1e-6)
    assert abs(p1[1] - p0[1]) < 1e-9  # orbit about the vertical axis
    assert (opt.width, opt.height, opt.ao_samples, opt.shadows) == (512, 512, 16, True)
    assert sc.camera.position == tuple(p0)  # the pool's scene is untouched


class _FakeServer(serve.Server):
    def __init__(self, budget, max_batch, resident):
        self.budget, self.tuning, self._resident = budget, serve.Tuning(max_batch=max_batch), resident

    def scene_bytes(self):
        return self._resident


def test_admission_by_hbm_budget():
    reqs = [serve.Request(i, 0.0, 500, "spheres", w, h, 0.0) for i, (w, h) in
            enumerate([(1920, 1080)] * 6 + [(512, 512)] * 4)]
    fb = serve.frame_bytes(1920, 1080)
    srv = _FakeServer(budget=3 * fb + 1000, max_batch=64, resident=1000)
    assert len(srv.admit(reqs)) == 3                      # the budget
    srv = _FakeServer(budget=1 << 40, max_batch=4, resident=0)
    assert len(srv.admit(reqs)) == 4                      # the batch size
    srv = _FakeServer(budget=1, max_batch=4, resident=0)
    assert len(srv.admit(reqs)) == 1                      # always at least one request
