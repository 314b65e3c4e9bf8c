"""The learned GI denoiser over the C-ABI (``ubg_unet_*``): a U-Net over the
irradiance field E and its guides that stands in for denoisePt1E / OIDN
(src/umbreon/integrator/pt1/pt1_denoise.cpp:20-81), trained on the device.

Features per pixel (NHWC float32, 12 channels): log(1 + E) rgb, albedo rgb,
normal xyz, depth / max depth, and two noise descriptors (the path tracer's
1 / sqrt(spp) and colour standard error); the GI gather's 10 features are
passed with the last two zero. The network predicts log(1 + E).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import capi

N_FEATURES = 12
GI_FEATURES = 10  # the GI gather's features (no noise descriptors)


def network_features(features: np.ndarray) -> np.ndarray:
    """(..., 12) float32 network input: 12-channel features as they are,
    the GI gather's 10 with two zero noise descriptors appended."""
    f = np.asarray(features, np.float32)
    if f.shape[-1] == GI_FEATURES:
        f = np.concatenate([f, np.zeros(f.shape[:-1] + (N_FEATURES - GI_FEATURES,), np.float32)], axis=-1)
    if f.shape[-1] != N_FEATURES:
        raise ValueError(f"expected {GI_FEATURES} or {N_FEATURES} features per pixel, got {f.shape[-1]}")
    return np.ascontiguousarray(f)


class Unet:
    """One U-Net on a device context (umbreon_amd.render.Device)."""

    def __init__(self, device, levels: int = 4, channels: Sequence[int] = (32, 64, 96, 128), seed: int = 0):
        self._lib = capi.load()
        self.device = device
        self.levels = int(levels)
        self.channels = tuple(int(c) for c in channels[:levels])
        ch = (C.c_int32 * 6)(*(list(self.channels) + [0] * (6 - len(self.channels))))
        h = C.c_void_p()
        capi.check(self._lib.ubg_unet_create(device.handle, self.levels, ch, C.c_uint64(seed), C.byref(h)))
        self.handle = h
        self.n_params = int(self._lib.ubg_unet_param_count(h))

    def close(self) -> None:
        if self.handle:
            self._lib.ubg_unet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_min_workgroups(self, n: int) -> None:
        """Convolution launch-shape threshold (test hook): 0 makes every
        convolution use the widest pixel tiles its rows allow."""
        capi.check(self._lib.ubgx_unet_set_min_workgroups(self.handle, int(n)))

    def set_loss(self, mode: str) -> None:
        """Training loss: 'l1' (mean absolute error, the default) or 'l2'
        (mean squared error)."""
        capi.check(self._lib.ubg_unet_set_loss(self.handle, {"l1": 0, "l2": 1}[mode]))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        capi.check(self._lib.ubg_unet_get_params(self.handle, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        capi.check(self._lib.ubg_unet_set_params(self.handle, p.ctypes.data))

    def infer(self, features: np.ndarray) -> np.ndarray:
        """features (n, h, w, 12 or 10) -> predicted log(1 + E) (n, h, w, 3)."""
        f = network_features(features)
        n, h, w, _ = f.shape
        out = np.empty((n, h, w, 3), np.float32)
        capi.check(self._lib.ubg_unet_infer(self.handle, n, h, w, f.ctypes.data, out.ctypes.data, 0))
        return out

    def train_step(self, features: np.ndarray, targets: np.ndarray, lr: float,
                   want_grads: bool = False) -> Tuple[float, Optional[np.ndarray]]:
        f = network_features(features)
        t = np.ascontiguousarray(targets, np.float32)
        n, h, w, _ = f.shape
        loss = C.c_float(0.0)
        g = np.empty(self.n_params, np.float32) if want_grads else None
        capi.check(self._lib.ubg_unet_train_step(self.handle, n, h, w, f.ctypes.data, t.ctypes.data, float(lr),
                                                  C.byref(loss), g.ctypes.data if g is not None else None, 0))
        return float(loss.value), g
