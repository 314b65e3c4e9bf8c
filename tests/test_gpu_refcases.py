"""The reference's closed-form render cases on the gfx950 path, and per-pixel
parity of every case with the CPU oracle (same inputs, same seeds).

Tolerance: the case's `tol` (1e-4 max abs on linear HDR for most cases; the
FP32 shading uses ocml pow/acos/cos/sin on the GPU vs glibc on the host, a few
ulp apart). Hit ids must agree exactly on these tiny scenes.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import binding as ob
from refcases import CASES

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def _gpu(device, scene, opt, **kw):
    from umbreon_amd.render import BuiltScene
    bs = BuiltScene(device, scene)
    try:
        return bs.render(scene, opt, want_ids=True, want_ao=True, **kw)
    finally:
        bs.close()


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_case_on_gpu(device, case):
    print(f"case {case.name}", flush=True)
    res = {k: _gpu(device, s, o) for k, (s, o) in case.renders.items()}
    case.check(res)


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_case_matches_oracle(device, case):
    print(f"case {case.name}", flush=True)
    for key, (s, o) in case.renders.items():
        g = _gpu(device, s, o, hires_only=True)
        c = ob.render(s, o, hires_only=True, threads=2)
        assert np.array_equal(g.hit_id, c.hit_id), f"{key}: first-hit ids differ"
        np.testing.assert_allclose(g.depth, c.depth, rtol=0, atol=1e-5 * max(1.0, float(np.abs(c.depth).max())))
        d = np.abs(g.color - c.color).max()
        assert d <= case.tol, f"{key}: max abs {d} > {case.tol}"
        for k in ("primary_rays", "shaded_hits", "shadow_rays", "capped_rays"):
            assert g.stats[k] == c.stats[k], (key, k, g.stats[k], c.stats[k])
        assert (g.stats["ao_rays_traced"] + g.stats["ao_rays_sampled"]
                == c.stats["ao_rays_traced"] + c.stats["ao_rays_sampled"])
