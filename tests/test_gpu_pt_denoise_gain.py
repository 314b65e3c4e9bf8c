"""The shipped path-traced denoiser (umbreon_amd/weights/pt_unet.npz) on
molecules it was never trained on: 256x256 windows of held-out seeds (the
training windows' framing: the full frame's pixel footprint), PSNR (sRGB,
whole window) of the plain and the denoised path tracer against a 1024-spp
render of the same window. The network must add at least 5 dB at 4 spp and
3 dB at 64 spp on every window (measured on gfx950: +9.1 / +8.0 dB at 4 spp,
+6.4 / +5.8 dB at 64 spp)."""
from __future__ import annotations

from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from umbreon_amd import capi

pytestmark = pytest.mark.gpu

WEIGHTS = Path(__file__).resolve().parent.parent / "umbreon_amd" / "weights" / "pt_unet.npz"


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("seed", [900000, 900001])
def test_learned_beats_plain_on_held_out_windows(device, seed):
    from umbreon_amd.denoise_train import load
    from umbreon_amd.image_io import psnr_srgb
    from umbreon_amd.pathtrace import pathtrace
    from umbreon_amd.pt_train import window_scene
    from umbreon_amd.render import BuiltScene
    sc, opt = window_scene(seed, 256)
    net = load(device, str(WEIGHTS))
    bs = BuiltScene(device, sc, capi.BVH8_GPU_BUILD)
    try:
        ref = pathtrace(bs, sc, replace(opt, spp=1024, seed=99)).color
        gains = {}
        for n in (4, 64):
            o = replace(opt, spp=n, seed=5)
            plain = psnr_srgb(pathtrace(bs, sc, o).color, ref)
            den = psnr_srgb(pathtrace(bs, sc, o, denoise_net=net).color, ref)
            gains[n] = den - plain
            print(f"held-out window {seed}, {n} spp: plain {plain:.2f} dB, learned {den:.2f} dB", flush=True)
    finally:
        bs.close()
        net.close()
    assert np.isfinite(list(gains.values())).all()
    assert gains[4] >= 5.0 and gains[64] >= 3.0, gains
