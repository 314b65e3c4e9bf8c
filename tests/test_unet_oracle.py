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


@pytest.mark.parametrize("loss", ["l1", "l2"])
@pytest.mark.parametrize("levels,ch,shape", [(2, [32, 32], (1, 4, 4)), (3, [32, 64, 32], (2, 8, 4))])
def test_gradients_match_finite_differences(levels, ch, shape, loss):
    rng = np.random.default_rng(levels)
    p = _params(levels, ch, rng)
    n, h, w = shape
    feat = rng.uniform(0.0, 1.0, (n, h, w, U.N_FEATURES))
    tgt = rng.uniform(0.0, 1.0, (n, h, w, 3)) + 5.0  # far from the prediction: a smooth L1 branch
    l0, g = U.loss_and_grads(p, feat, tgt, levels, ch, bf16=False, loss=loss)
    pred, _ = U.forward(p, feat, levels, ch, bf16=False)
    d = pred - tgt
    assert l0 == pytest.approx(float((d * d).mean() if loss == "l2" else np.abs(d).mean()), rel=1e-12)
    idx = rng.choice(np.flatnonzero(g != 0.0), 40, replace=False)
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = U.loss_and_grads(q, feat, tgt, levels, ch, bf16=False, loss=loss)
        q[i] -= 2 * eps
        lm, _ = U.loss_and_grads(q, feat, tgt, levels, ch, bf16=False, loss=loss)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), abs(g[i])) + 1e-12, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.1415926], np.float32)
    r = U.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


# ---------------------------------------------------------------------------
# The forward pass against an independent statement of the same network in
# torch.nn.functional (float64, CPU): conv2d with padding 1 (cross-correlation,
# weight[cout, cin, ky, kx] = W[3 ky + kx][cout][cin]), relu, avg_pool2d(2),
# nearest upsampling by 2 and channel concatenation [skip, upsampled]. The
# architecture is written out here from its description -- per level two
# convolutions (the first reads the previous level pooled), per decoder level
# a convolution of [encoder output, upsampled deeper output] and a second one,
# then the output convolution whose channels 0..2 are added to log(1 + E) of
# the input -- and the parameters are walked in that order, so a wrong tap
# orientation, pooling window, upsampling index, concat order or parameter
# offset in oracle/unet.py fails it.

def _torch_forward(params, feat, levels, ch):
    import torch
    import torch.nn.functional as F
    p = torch.from_numpy(np.asarray(params, np.float64))
    off = [0]

    def conv(x, cout, relu):
        cin = x.shape[1]
        nw = 9 * cout * cin
        w = p[off[0]:off[0] + nw].reshape(3, 3, cout, cin).permute(2, 3, 0, 1)
        b = p[off[0] + nw:off[0] + nw + cout]
        off[0] += nw + cout
        y = F.conv2d(x, w, b, padding=1)
        return F.relu(y) if relu else y

    f = torch.from_numpy(np.asarray(feat, np.float64)).permute(0, 3, 1, 2)
    x = torch.zeros((f.shape[0], U.IN_PAD, f.shape[2], f.shape[3]), dtype=torch.float64)
    x[:, :U.N_FEATURES] = f
    enc = []
    for lv in range(levels):
        if lv > 0:
            x = F.avg_pool2d(x, 2)
        x = conv(x, ch[lv], True)
        x = conv(x, ch[lv], True)
        enc.append(x)
    d = enc[-1]
    for lv in range(levels - 2, -1, -1):
        up = F.interpolate(d, scale_factor=2, mode="nearest")
        d = conv(torch.cat([enc[lv], up], dim=1), ch[lv], True)
        d = conv(d, ch[lv], True)
    out = conv(d, U.OUT_PAD, False)
    assert off[0] == U.param_count(levels, ch)
    return (out[:, :3] + f[:, :3]).permute(0, 2, 3, 1).numpy()


@pytest.mark.parametrize("levels,ch,shape", [(2, [32, 64], (2, 8, 12)), (3, [32, 64, 96], (1, 16, 8))])
def test_forward_matches_torch(levels, ch, shape):
    rng = np.random.default_rng(11 + levels)
    p = rng.normal(0.0, 0.15, U.param_count(levels, ch))
    n, h, w = shape
    feat = rng.uniform(-1.0, 1.0, (n, h, w, U.N_FEATURES))
    got, _ = U.forward(p, feat, levels, ch, bf16=False)
    want = _torch_forward(p, feat, levels, ch)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, scale)


def test_conv_delta_kernel_is_a_shift():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 6, 7, 4))
    for t, (dy, dx) in enumerate([(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 0), (0, 1), (1, -1), (1, 0), (1, 1)]):
        w = np.zeros((9, 4, 4))
        w[t] = np.eye(4)
        y = U._conv(x, w)
        want = np.zeros_like(x)
        for yy in range(6):
            for xx in range(7):
                sy, sx = yy + dy, xx + dx
                if 0 <= sy < 6 and 0 <= sx < 7:
                    want[:, yy, xx] = x[:, sy, sx]
        assert np.array_equal(y, want), t


def test_conv_box_kernel_on_a_constant_field():
    # every tap weight 1 between all channel pairs: interior = 9 * cin * c,
    # edges 6 taps, corners 4 (zero padding)
    c, cin = 0.75, 3
    x = np.full((1, 5, 6, cin), c)
    y = U._conv(x, np.ones((9, 2, cin)))
    taps = np.full((5, 6), 9.0)
    taps[0, :] = taps[-1, :] = 6.0
    taps[:, 0] = taps[:, -1] = 6.0
    taps[0, 0] = taps[0, -1] = taps[-1, 0] = taps[-1, -1] = 4.0
    assert np.array_equal(y[0, :, :, 0], taps * cin * c) and np.array_equal(y[0, :, :, 1], taps * cin * c)


def test_pool_and_upsample_closed_forms():
    x = np.arange(2 * 4 * 6 * 1, dtype=np.float64).reshape(2, 4, 6, 1)
    pooled = U._pool(x, bf16=False)
    assert pooled.shape == (2, 2, 3, 1)
    assert pooled[0, 0, 0, 0] == (0 + 1 + 6 + 7) / 4 and pooled[1, 1, 2, 0] == (40 + 41 + 46 + 47) / 4
    acts = {0: x}
    up = U._operand(acts, (0, 2), bf16=False)
    assert up.shape == (2, 8, 12, 1)
    assert up[0, 3, 5, 0] == x[0, 1, 2, 0] and up[1, 7, 11, 0] == x[1, 3, 5, 0]


def test_network_features_pads_the_gather_features():
    # the GI gather writes 10 features; the network reads 12, the two noise
    # descriptors zero for the gather, and 12-channel features pass unchanged
    from umbreon_amd.denoise import N_FEATURES, network_features
    rng = np.random.default_rng(3)
    f10 = rng.uniform(-1.0, 1.0, (2, 4, 8, 10)).astype(np.float32)
    f = network_features(f10)
    assert f.shape == (2, 4, 8, N_FEATURES) and f.dtype == np.float32 and f.flags.c_contiguous
    assert np.array_equal(f[..., :10], f10) and not f[..., 10:].any()
    f12 = rng.uniform(-1.0, 1.0, (1, 4, 8, 12)).astype(np.float32)
    assert np.array_equal(network_features(f12), f12)
    with pytest.raises(ValueError):
        network_features(np.zeros((1, 4, 8, 11), np.float32))
