"""The learned denoiser's device U-Net against the oracle (oracle/unet.py):
forward predictions and parameter gradients on the same bf16 operands, and
that training lowers the loss."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import unet as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def _batch(rng, n, h, w):
    feat = np.zeros((n, h, w, 12), np.float32)
    feat[..., 0:3] = rng.uniform(0.0, 2.0, (n, h, w, 3))
    feat[..., 3:6] = rng.uniform(0.0, 1.0, (n, h, w, 3))
    nrm = rng.normal(size=(n, h, w, 3))
    feat[..., 6:9] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    feat[..., 9] = rng.uniform(0.0, 1.0, (n, h, w))
    feat[..., 10] = rng.uniform(1.0 / 64, 1.0, (n, h, w))  # 1 / sqrt(spp)
    feat[..., 11] = rng.uniform(0.0, 0.5, (n, h, w))  # colour standard error
    tgt = (feat[..., 0:3] * 0.5 + 0.2).astype(np.float32)
    return feat, tgt


# min_wgs 0 makes every convolution take the widest pixel tiles its rows
# allow, so these small images run the instantiations that 512^2 and 1024^2
# frames use: (1, 16, 256) with 32/64/96/128 channels covers NT 2 / NP 4 at
# level 0, NT 4 / NP 2, NT 6 / NP 2 and NT 8 / NP 2 below it, forward and
# (transposed) backward.
@pytest.mark.parametrize("levels,ch,shape,min_wgs", [
    (2, [32, 64], (2, 16, 32), 1024),
    (4, [32, 64, 96, 128], (1, 32, 48), 1024),
    (4, [32, 64, 96, 128], (1, 16, 256), 0),
    (3, [32, 64, 96], (2, 16, 128), 0),
])
def test_forward_and_gradients_match_oracle(device, levels, ch, shape, min_wgs):
    from umbreon_amd.denoise import Unet
    net = Unet(device, levels, ch, seed=3)
    net.set_min_workgroups(min_wgs)
    try:
        assert net.n_params == U.param_count(levels, ch)
        p = net.get_params()
        rng = np.random.default_rng(7)
        # non-zero biases so every term is exercised
        _, convs = U.layout(levels, ch)
        for c in convs:
            p[c["b_off"]:c["b_off"] + c["cout"]] = rng.normal(0.0, 0.05, c["cout"])
        net.set_params(p)
        feat, tgt = _batch(rng, *shape)
        g_pred = net.infer(feat)
        c_pred, _ = U.forward(p.astype(np.float64), feat, levels, ch, bf16=True)
        d = np.abs(g_pred - c_pred)
        print(f"unet L{levels} {shape} min_wgs {min_wgs} forward: max|d| {d.max():.3g} mean|d| {d.mean():.3g}",
              flush=True)
        assert d.mean() <= 2e-3 and d.max() <= 5e-2
        g_loss, g_grads = net.train_step(feat, tgt, lr=0.0, want_grads=True)
        c_loss, c_grads = U.loss_and_grads(p.astype(np.float64), feat, tgt, levels, ch, bf16=True)
        assert abs(g_loss - c_loss) <= 1e-3 * c_loss
        for k, c in enumerate(convs):
            a = g_grads[c["w_off"]:c["b_off"] + c["cout"]].astype(np.float64)
            b = c_grads[c["w_off"]:c["b_off"] + c["cout"]]
            nb = np.linalg.norm(b)
            if nb == 0.0:
                assert np.linalg.norm(a) == 0.0
                continue
            cos = float(a @ b / (np.linalg.norm(a) * nb))
            rel = float(np.linalg.norm(a - b) / nb)
            print(f"  conv {k} ({c['cin']}->{c['cout']}): cos {cos:.6f} rel {rel:.3g}", flush=True)
            assert cos >= 0.999 and rel <= 5e-2, (k, cos, rel)
    finally:
        net.close()


def test_squared_loss_gradients_match_oracle(device):
    """The mean-squared training loss (ubg_unet_set_loss 1) on the device
    against oracle/unet.py's (its gradients pinned by finite differences)."""
    from umbreon_amd.denoise import Unet
    levels, ch = 3, [32, 64, 96]
    net = Unet(device, levels, ch, seed=5)
    try:
        net.set_loss("l2")
        p = net.get_params()
        rng = np.random.default_rng(17)
        _, convs = U.layout(levels, ch)
        for c in convs:
            p[c["b_off"]:c["b_off"] + c["cout"]] = rng.normal(0.0, 0.05, c["cout"])
        net.set_params(p)
        feat, tgt = _batch(rng, 2, 16, 32)
        g_loss, g_grads = net.train_step(feat, tgt, lr=0.0, want_grads=True)
        c_loss, c_grads = U.loss_and_grads(p.astype(np.float64), feat, tgt, levels, ch, bf16=True, loss="l2")
        l1, _ = U.loss_and_grads(p.astype(np.float64), feat, tgt, levels, ch, bf16=True, loss="l1")
        print(f"unet squared loss: device {g_loss:.6g} oracle {c_loss:.6g} (L1 {l1:.6g})", flush=True)
        assert abs(g_loss - c_loss) <= 2e-3 * c_loss and abs(g_loss - l1) > 1e-2 * l1
        for k, c in enumerate(convs):
            a = g_grads[c["w_off"]:c["b_off"] + c["cout"]].astype(np.float64)
            b = c_grads[c["w_off"]:c["b_off"] + c["cout"]]
            nb = np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * nb))
            rel = float(np.linalg.norm(a - b) / nb)
            assert cos >= 0.999 and rel <= 5e-2, (k, cos, rel)
        net.set_loss("l1")
        g1, _ = net.train_step(feat, tgt, lr=0.0)
        assert abs(g1 - l1) <= 2e-3 * l1
    finally:
        net.close()


def test_forward_at_a_training_crop(device):
    """A 128^2 crop through the 4-level net at the default launch shapes.
    Device and oracle round the same operands to bf16, so they differ only
    by fp32 vs float64 accumulation and the bf16 roundings that difference
    flips on the way through 17 convolutions: bounded here by bf16 ulps
    (2^-8 relative) of the output's scale -- 4 of its mean on the mean error,
    8 of its max on the max error (measured on gfx950: 1.8 and 2.7)."""
    from umbreon_amd.denoise import Unet
    levels, ch = 4, [32, 64, 96, 128]
    net = Unet(device, levels, ch, seed=9)
    try:
        p = net.get_params()
        rng = np.random.default_rng(13)
        _, convs = U.layout(levels, ch)
        for c in convs:
            p[c["b_off"]:c["b_off"] + c["cout"]] = rng.normal(0.0, 0.05, c["cout"])
        net.set_params(p)
        feat, _ = _batch(rng, 1, 128, 128)
        g_pred = net.infer(feat)
        c_pred, cache = U.forward(p.astype(np.float64), feat, levels, ch, bf16=True)
        z = np.abs(cache["out"][..., :3])
        d = np.abs(g_pred - c_pred)
        ulp = 2.0 ** -8
        print(f"unet 128^2 forward: max|d| {d.max():.3g} mean|d| {d.mean():.3g} (output scale max {z.max():.3g} "
              f"mean {z.mean():.3g})", flush=True)
        assert d.mean() <= 4 * ulp * z.mean()
        assert d.max() <= 8 * ulp * z.max()
    finally:
        net.close()


def test_training_lowers_the_loss(device):
    from umbreon_amd.denoise import Unet
    net = Unet(device, 3, [32, 64, 96], seed=1)
    try:
        rng = np.random.default_rng(2)
        feat, tgt = _batch(rng, 4, 32, 32)
        first, _ = net.train_step(feat, tgt, lr=0.0)
        for _ in range(150):
            loss, _ = net.train_step(feat, tgt, lr=2e-3)
        print(f"unet training: loss {first:.4f} -> {loss:.4f}", flush=True)
        assert loss < 0.5 * first
    finally:
        net.close()


def test_gi_features_match_oracle(device):
    """The gi_features AOV (log(1 + E) of the gather, albedo, normal AOV,
    depth / max) against the oracle's, and the learned denoiser running in
    the render in place of the a-trous filter."""
    from umbreon_amd.denoise import Unet
    from umbreon_amd.denoise_train import gi_options, training_scene
    from umbreon_amd.render import BuiltScene
    from oracle import binding as ob
    sc = training_scene(1, 96, 64)
    opt = gi_options(96, 64, 4, seed=3)
    bs = BuiltScene(device, sc)
    net = Unet(device, 3, [32, 64, 96], seed=5)
    try:
        g = bs.render(sc, opt, hires_only=True, want_gi_features=True)
        c = ob.render(sc, opt, hires_only=True, want_gi_features=True)
        d = np.abs(g.gi_features - c.gi_features)
        print(f"gi features: max|d| per channel {np.round(d.max(axis=(0, 1)), 6)}", flush=True)
        assert np.array_equal(g.gi_features[..., 3:10], c.gi_features[..., 3:10])
        assert d[..., :3].max(axis=2).mean() <= 1e-3
        # the learned path: the receivers' E replaced by the network's
        # prediction, everything else as without a denoiser
        fl = bs.render(sc, opt, hires_only=True, denoise_net=net, want_gi_features=True)
        assert np.array_equal(fl.gi_features, g.gi_features)  # the features are taken before the denoise
        recv = g.gi_features[..., 3:6].max(axis=2) > 0
        assert np.array_equal(fl.color[~recv], g.color[~recv])
        assert not np.array_equal(fl.color[recv], g.color[recv])
        # pad + predict + apply equals the network on the padded features
        m = 4
        Hp, Wp = (64 + m - 1) // m * m, (96 + m - 1) // m * m
        pred = net.infer(np.pad(g.gi_features, ((0, Hp - 64), (0, Wp - 96), (0, 0)), mode="edge")[None])[0, :64, :96]
        E_net = np.maximum(np.expm1(pred), 0.0)
        refl = g.gi_features[..., 3:6]
        ind = fl.indirect
        assert np.allclose(ind[recv], (opt.gi_intensity * refl * E_net)[recv], rtol=1e-4, atol=1e-6)
    finally:
        net.close()
        bs.close()
