"""The gfx950 path against the committed FP64 golden fixtures (same gates as
the FP32 oracle must meet, tests/test_golden.py)."""
from __future__ import annotations

import pytest

from test_golden import NAMES, gates, load
import make_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("name", NAMES)
def test_gpu_meets_gates_against_fixture(device, name):
    from umbreon_amd.render import BuiltScene
    z, meta = load(name)
    sc, opt = make_golden.build(name)
    bs = BuiltScene(device, sc)
    try:
        g = bs.render(sc, opt, hires_only=True, want_ids=True, want_ao=True)
    finally:
        bs.close()
    gates(f"gpu {name}", g.color, g.hit_id, g.ao_open, g.stats, z, meta)
