"""What the trained denoiser gains over the method without it: on held-out
molecules (seeds the training never saw), the frame with the learned
denoiser at 2 samples per pixel against a 1024-sample render is several dB
closer than the plain path tracer and than the a-trous filter at the same
samples (umbreon_amd/weights/gi_unet.npz, trained by
``python -m umbreon_amd.denoise_train --scenes 192 --steps 6000``)."""
from __future__ import annotations

from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = Path(__file__).resolve().parent.parent / "umbreon_amd" / "weights" / "gi_unet.npz"


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("seed", [100000, 100001, 100002])
def test_learned_denoiser_gain(device, seed):
    from umbreon_amd.denoise_train import gi_options, load, training_scene
    from umbreon_amd.image_io import psnr_srgb
    from umbreon_amd.render import BuiltScene
    net = load(device, str(WEIGHTS))
    sc = training_scene(seed, 256, 256)
    bs = BuiltScene(device, sc)
    try:
        ref = bs.render(sc, gi_options(256, 256, 1024, seed=9000 + seed)).color
        noisy = gi_options(256, 256, 2, seed=100 + seed)
        plain = psnr_srgb(bs.render(sc, noisy).color, ref)
        atrous = psnr_srgb(bs.render(sc, replace(noisy, pt1_denoise=True)).color, ref)
        learned = psnr_srgb(bs.render(sc, noisy, denoise_net=net).color, ref)
    finally:
        bs.close()
        net.close()
    print(f"held-out scene {seed}, 2 spp vs 1024 spp: plain {plain:.2f} dB, a-trous {atrous:.2f} dB, "
          f"learned {learned:.2f} dB", flush=True)
    assert learned >= plain + 5.0
    assert learned >= atrous + 2.0
