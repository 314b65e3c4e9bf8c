"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle): the scene generator still makes the same scenes, the
FP64 oracle still renders them identically, and the FP32 oracle -- the
restatement the GPU is held to -- meets the parity gates against them."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as ob

GOLDEN = Path(__file__).resolve().parent / "golden"

import sys  # noqa: E402
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load(name):
    z = np.load(GOLDEN / f"{name}.npz")
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def gates(tag, color, hit_id, ao, stats, z, meta):
    ids = float((hit_id == z["hit_id"]).mean())
    d = np.abs(color - z["color"]).max(axis=2)
    bad = float((d > 1e-3).mean())
    aod = float(np.abs(ao - z["ao_open"]).mean())
    msg = f"{tag}: ids {ids:.6f}, max|d| {d.max():.3g}, >1e-3 {bad:.2e}, AO mean|d| {aod:.2e}"
    print(msg, flush=True)
    assert ids >= 0.9999 - 1.0 / hit_id.size, msg  # at most one pixel on these small frames
    assert bad <= 5e-4 + 1.0 / hit_id.size, msg
    assert aod <= 1e-3, msg
    for k in ("primary_rays", "shaded_hits", "shadow_rays", "capped_rays"):
        assert stats[k] == meta["stats"][k], (k, stats[k], meta["stats"][k])
    assert (stats["ao_rays_traced"] + stats["ao_rays_sampled"]
            == meta["stats"]["ao_rays_traced"] + meta["stats"]["ao_rays_sampled"])


@pytest.mark.parametrize("name", NAMES)
def test_scene_generator_unchanged(name):
    _, meta = load(name)
    sc, _ = make_golden.build(name)
    assert make_golden.scene_digest(sc) == meta["scene_sha256"]


@pytest.mark.parametrize("name", NAMES)
def test_fp64_oracle_reproduces_fixture(name):
    z, meta = load(name)
    sc, opt = make_golden.build(name)
    f = make_golden.render64(sc, opt)
    assert np.array_equal(f.hit_id, z["hit_id"])
    assert np.abs(f.color - z["color"]).max() <= 1e-9
    assert np.abs(f.ao_open - z["ao_open"]).max() <= 1e-12
    # (counters added after the fixture was written, e.g. the GI rays, are 0 here)
    assert {k: f.stats[k] for k in meta["stats"]} == meta["stats"]
    assert all(f.stats[k] == 0 for k in f.stats if k not in meta["stats"])


@pytest.mark.parametrize("name", NAMES)
def test_fp32_oracle_meets_gates_against_fixture(name):
    z, meta = load(name)
    sc, opt = make_golden.build(name)
    f = ob.render(sc, opt, hires_only=True)
    gates(f"fp32 oracle {name}", f.color, f.hit_id, f.ao_open, f.stats, z, meta)


def test_psnr_helper_closed_forms():
    from umbreon_amd.image_io import linear_to_srgb, psnr_srgb
    assert abs(float(linear_to_srgb(np.array(0.5))) - 0.7353569830524495) < 1e-12  # 188/255 in the reference's test
    a = np.zeros((2, 2, 4))
    b = a.copy()
    assert psnr_srgb(a, b) == 99.0
    b[0, 0, 0] = 1.0  # one channel of 12 off by 1.0 in sRGB
    assert abs(psnr_srgb(a, b) - 10.0 * np.log10(12.0)) < 1e-12
