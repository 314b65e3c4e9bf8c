"""CPU restatement of the learned denoiser's U-Net -- TEST INFRASTRUCTURE.

numpy forward and backward of the same network, parameter layout and loss
as the device implementation (umbreon_amd/csrc/ub_unet.hip): per
convolution W[9][cout][cin] then b[cout]; 3x3 convolutions with zero
padding; ReLU; 2x2 average pooling and nearest upsampling folded into the
next convolution's operand; L1 loss on log(1 + E) with the output added to
the input's log(1 + E). With ``bf16=True`` the operands are rounded to
bfloat16 where the device rounds them (inputs, weights, activations, the
masked output gradient), everything else is float64; with ``bf16=False`` the
whole network is float64 (the form the finite-difference checks pin).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

IN_PAD = 32
OUT_PAD = 32
N_FEATURES = 12  # log(1 + E) rgb, albedo rgb, normal xyz, depth, two noise descriptors


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(levels: int, channels) -> Tuple[list, list]:
    """(activations [(C, level)], convolutions [dict]) in forward order, the
    device's construction."""
    acts = [(IN_PAD, 0)]
    convs = []

    def act(c, lv):
        acts.append((c, lv))
        return len(acts) - 1

    def conv(lv, s0, s1, cout, relu, out):
        cin = acts[s0[0]][0] + (acts[s1[0]][0] if s1 else 0)
        convs.append(dict(level=lv, s0=s0, s1=s1, cin=cin, cout=cout, relu=relu, out=out))

    e = []
    prev = 0
    for lv in range(levels):
        a = act(channels[lv], lv)
        conv(lv, (prev, 0 if lv == 0 else 1), None, channels[lv], 1, a)
        e.append(act(channels[lv], lv))
        conv(lv, (a, 0), None, channels[lv], 1, e[lv])
        prev = e[lv]
    d = e[-1]
    for lv in range(levels - 2, -1, -1):
        a = act(channels[lv], lv)
        conv(lv, (e[lv], 0), (d, 2), channels[lv], 1, a)
        b = act(channels[lv], lv)
        conv(lv, (a, 0), None, channels[lv], 1, b)
        d = b
    conv(0, (d, 0), None, OUT_PAD, 0, -1)
    off = 0
    for c in convs:
        c["w_off"] = off
        off += 9 * c["cout"] * c["cin"]
        c["b_off"] = off
        off += c["cout"]
    return acts, convs


def param_count(levels: int, channels) -> int:
    _, convs = layout(levels, channels)
    return convs[-1]["b_off"] + convs[-1]["cout"]


def _taps():
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """out[n, y, x] = x[n, y + dy, x + dx], zero outside."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, ye = max(0, -dy), min(h, h - dy)
    xs, xe = max(0, -dx), min(w, w - dx)
    out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def _pool(x: np.ndarray, bf16: bool) -> np.ndarray:
    if bf16:
        f = x.astype(np.float32)
        s = ((f[:, 0::2, 0::2] + f[:, 0::2, 1::2]) + f[:, 1::2, 0::2]) + f[:, 1::2, 1::2]
        return to_bf16(s * np.float32(0.25))
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def _operand(acts_v, src, bf16):
    a, mode = src
    x = acts_v[a]
    if mode == 1:
        return _pool(x, bf16)
    if mode == 2:
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)
    return x


def _conv(x: np.ndarray, w: np.ndarray) -> np.ndarray:
    """x (n, h, w, cin), w (9, cout, cin) -> (n, h, w, cout)."""
    out = 0.0
    for t, (dy, dx) in enumerate(_taps()):
        out = out + _shift(x, dy, dx) @ w[t].T
    return out


def forward(params: np.ndarray, feat: np.ndarray, levels: int, channels, bf16: bool = True):
    """Predicted log(1 + E) (n, h, w, 3) and the cache for backward."""
    acts, convs = layout(levels, channels)
    rnd = to_bf16 if bf16 else (lambda v: np.asarray(v, np.float64))
    n, h, w, nf = feat.shape
    assert nf == N_FEATURES, nf
    x0 = np.zeros((n, h, w, IN_PAD))
    x0[..., :N_FEATURES] = rnd(feat)
    vals = {0: x0}
    xts, out = [], None
    for c in convs:
        xt = _operand(vals, c["s0"], bf16)
        if c["s1"]:
            xt = np.concatenate([xt, _operand(vals, c["s1"], bf16)], axis=-1)
        wgt = rnd(params[c["w_off"]:c["b_off"]].reshape(9, c["cout"], c["cin"]))
        b = params[c["b_off"]:c["b_off"] + c["cout"]].astype(np.float64)
        z = _conv(xt, wgt) + b
        if bf16:
            z = z.astype(np.float32).astype(np.float64)
        xts.append(xt)
        if c["relu"]:
            z = np.maximum(z, 0.0)
        if c["out"] >= 0:
            vals[c["out"]] = rnd(z)
        else:
            out = z
    pred = out[..., :3] + np.asarray(feat[..., :3], np.float64)
    return pred, dict(vals=vals, xts=xts, out=out)


def loss_and_grads(params: np.ndarray, feat: np.ndarray, target: np.ndarray, levels: int, channels,
                   bf16: bool = True, loss: str = "l1"):
    """Mean L1 (loss 'l1') or mean squared (loss 'l2') loss on the
    prediction and the parameter gradients (ub_unet.hip k_unet_loss)."""
    acts, convs = layout(levels, channels)
    rnd = to_bf16 if bf16 else (lambda v: np.asarray(v, np.float64))
    pred, cache = forward(params, feat, levels, channels, bf16)
    d = pred - np.asarray(target, np.float64)
    count = d.size
    l2 = loss == "l2"
    loss = float((d * d).mean()) if l2 else float(np.abs(d).mean())
    grads = np.zeros(params.size, np.float64)
    n, h, w, _ = feat.shape
    dout = np.zeros(cache["out"].shape)
    dout[..., :3] = (2.0 * d if l2 else np.sign(d)) / count
    dacts = {a: np.zeros(cache["vals"][a].shape) for a in cache["vals"] if a > 0}
    for k in range(len(convs) - 1, -1, -1):
        c = convs[k]
        dy = dout if c["out"] < 0 else dacts[c["out"]]
        if c["out"] >= 0 and c["relu"]:
            dy = dy * (cache["vals"][c["out"]] > 0)
        grads[c["b_off"]:c["b_off"] + c["cout"]] = dy.sum(axis=(0, 1, 2))
        dz = rnd(dy)
        xt = cache["xts"][k]
        gw = np.zeros((9, c["cout"], c["cin"]))
        for t, (ty, tx) in enumerate(_taps()):
            gw[t] = np.einsum("nyxo,nyxi->oi", dz, _shift(xt, ty, tx))
        grads[c["w_off"]:c["b_off"]] = gw.reshape(-1)
        need0 = c["s0"][0] > 0
        need1 = bool(c["s1"]) and c["s1"][0] > 0
        if not (need0 or need1):
            continue
        wgt = rnd(params[c["w_off"]:c["b_off"]].reshape(9, c["cout"], c["cin"]))
        dx = 0.0
        for t, (ty, tx) in enumerate(_taps()):
            dx = dx + _shift(dz, -ty, -tx) @ wgt[t]
        if bf16:
            dx = dx.astype(np.float32).astype(np.float64)

        def scatter(src, off):
            a, mode = src
            cs = acts[a][0]
            g = dx[..., off:off + cs]
            if mode == 1:
                g = 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)
            elif mode == 2:
                g = g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]
            dacts[a] += g

        if need0:
            scatter(c["s0"], 0)
        if need1:
            scatter(c["s1"], acts[c["s0"][0]][0])
    return loss, grads
