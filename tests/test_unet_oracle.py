"""Pins for the oracle's U-Net (oracle/unet.py) -- CPU only: its analytic
gradients against central finite differences of its own float64 forward
(a wrong transpose, tap flip, pooling / upsampling adjoint, concat offset or
bias term fails them), and the layout against hand counts."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet as U


def _params(levels, ch, rng):
    p = rng.normal(0.0, 0.2, U.param_count(levels, ch))
    _, convs = U.layout(levels, ch)
    first, last = convs[0], convs[-1]
    p[first["w_off"]:first["b_off"]].reshape(9, first["cout"], first["cin"])[:, :, U.N_FEATURES:] = 0.0
    p[last["w_off"]:last["b_off"]].reshape(9, last["cout"], last["cin"])[:, 3:, :] = 0.0
    return p


def test_layout_counts():
    # 2 levels of 32: enc 32->32, 32->32, 32->32, 32->32; dec (32+32)->32,
    # 32->32; out 32->32: 9 cin cout + cout each
    convs = [(32, 32)] * 4 + [(64, 32), (32, 32), (32, 32)]
    assert U.param_count(2, [32, 32]) == sum(9 * ci * co + co for ci, co in convs)
    acts, cv = U.layout(3, [32, 64, 96])
    assert len(cv) == 2 * 3 + 2 * 2 + 1
    assert [c["s1"][1] for c in cv if c["s1"]] == [2, 2]  # decoder inputs upsampled
    assert cv[2]["s0"][1] == 1 and cv[4]["s0"][1] == 1  # encoder inputs pooled


@pytest.mark.parametrize("levels,ch,shape", [(2, [32, 32], (1, 4, 4)), (3, [32, 64, 32], (2, 8, 4))])
def test_gradients_match_finite_differences(levels, ch, shape):
    rng = np.random.default_rng(levels)
    p = _params(levels, ch, rng)
    n, h, w = shape
    feat = rng.uniform(0.0, 1.0, (n, h, w, U.N_FEATURES))
    tgt = rng.uniform(0.0, 1.0, (n, h, w, 3)) + 5.0  # far from the prediction: a smooth L1 branch
    _, g = U.loss_and_grads(p, feat, tgt, levels, ch, bf16=False)
    idx = rng.choice(np.flatnonzero(g != 0.0), 40, replace=False)
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = U.loss_and_grads(q, feat, tgt, levels, ch, bf16=False)
        q[i] -= 2 * eps
        lm, _ = U.loss_and_grads(q, feat, tgt, levels, ch, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), abs(g[i])) + 1e-12, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.1415926], np.float32)
    r = U.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
