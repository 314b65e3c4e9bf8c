"""Image comparison as the reference's bench harness does it
(src/bench/image/image_io.cpp:16-20, 223-238): PSNR over sRGB-encoded RGB,
each linear channel clamped to [0, 1] first; 99 dB for identical images."""
from __future__ import annotations

import numpy as np


def linear_to_srgb(c: np.ndarray) -> np.ndarray:
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1.0 / 2.4) - 0.055)


def psnr_srgb(a: np.ndarray, b: np.ndarray) -> float:
    """a, b: (H, W, >=3) linear RGB(A); alpha is ignored."""
    d = linear_to_srgb(a[..., :3]) - linear_to_srgb(b[..., :3])
    mse = float(np.mean(d * d))
    return 99.0 if mse <= 1e-12 else 10.0 * np.log10(1.0 / mse)
