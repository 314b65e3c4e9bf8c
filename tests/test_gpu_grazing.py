"""The FP64 root polish on the GPU: the grazing rays of tests/grazing.py
(spheres, capped cones, round curves and their end spheres ~1e4 from the
origin, passing 1e-6 .. 3e-3 of a radius inside or outside the silhouette)
through ubg_intersect / ubg_occluded. The device classifies every ray as the
FP64 oracle does and returns the library's host-side results (the same
intersectors compiled for the CPU, equal to the FP32 oracle) bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from grazing import curve_case, sphere_case
from oracle import binding as ob
from umbreon_amd import capi
from umbreon_amd.render import host_query

pytestmark = pytest.mark.gpu

CASES = {
    "sphere": lambda rng: sphere_case(rng),
    "cone": lambda rng: curve_case(rng, open_=False),
    "round": lambda rng: curve_case(rng, open_=True),
    "round_end": lambda rng: curve_case(rng, open_=True, at_end=True),
}


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_grazing_rays(device, name):
    from umbreon_amd.render import BuiltScene
    rng = np.random.default_rng(17)
    sc, r, _ = CASES[name](rng)
    truth = ob.intersect(sc, r, precision=64, brute=True)
    th = truth["kind"] != 0xFFFFFFFF
    host, _, _ = host_query(sc, r, capi.BVH8)
    bs = BuiltScene(device, sc)
    try:
        g = bs.intersect(r)
        occ = bs.occluded(r)
    finally:
        bs.close()
    gh = g["kind"] != 0xFFFFFFFF
    print(f"{name}: {int(th.sum())} of {len(r)} rays hit; GPU misclassified {int((gh != th).sum())}", flush=True)
    assert np.array_equal(gh, th)
    assert np.array_equal(occ.astype(bool), th)
    assert np.array_equal(g["t"].view(np.uint32), host["t"].view(np.uint32))
    assert np.array_equal(g["prim_id"], host["prim_id"]) and np.array_equal(g["kind"], host["kind"])
