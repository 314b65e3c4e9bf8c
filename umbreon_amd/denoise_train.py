"""Training and evaluation of the learned GI denoiser on pairs the renderer
generates itself: seeded synthetic molecules (CPK spheres, ball-and-stick,
molecular surfaces) rendered with pt2 GI at 1-8 samples per pixel (the
noisy input features) and at a high sample count (the target irradiance).

    python -m umbreon_amd.denoise_train --scenes 192 --steps 6000 --out umbreon_amd/weights/gi_unet.npz
    python -m umbreon_amd.denoise_train --evaluate umbreon_amd/weights/gi_unet.npz
    python -m umbreon_amd.denoise_train --bench-infer 1024 --calls 11

The evaluation reports, per held-out scene and for the full frame (linear
colour, sRGB PSNR against a 4096-sample render as the reference's bench
compares images), the frame time and PSNR of the plain path tracer at
increasing sample counts and of the learned and a-trous denoisers at low
counts, and the time each needs to reach 40 dB.
"""
from __future__ import annotations

import argparse
import math
import time
from dataclasses import replace
from typing import List, Optional, Tuple

import numpy as np

from .denoise import Unet
from .image_io import psnr_srgb
from .render import BuiltScene, Device
from .scene import RenderOptions, Scene
from .scenes import ball_and_stick_scene, molecular_surface_scene, spheres_scene

def _say(msg: str) -> None:
    print(msg, flush=True)


DEFAULT_LEVELS = 4
DEFAULT_CHANNELS = (32, 64, 96, 128)


def gi_options(width: int, height: int, spp: int, seed: int) -> RenderOptions:
    """pt2 GI (the reference default integrator, blue-noise first bounce) on
    the full-resolution gather, no denoise, hard shadows."""
    return RenderOptions(width=width, height=height, supersample=1, shadows=True, ao_samples=0, gi=True,
                         gi_integrator=2, gi_bounces=2, pt1_spp=spp, pt1_seed=seed, pt1_gather_div=1,
                         pt1_denoise=False, pt1_edge_patch=False)


def training_scene(seed: int, width: int, height: int) -> Scene:
    """One seeded synthetic molecule: the kind, size and light vary with the
    seed."""
    rng = np.random.default_rng(seed)
    kind = seed % 3
    if kind == 0:
        sc, _ = spheres_scene(n_atoms=int(rng.integers(300, 2500)), width=width, height=height, seed=seed)
    elif kind == 1:
        sc, _ = ball_and_stick_scene(n_atoms=int(rng.integers(300, 3000)), width=width, height=height, seed=seed)
    else:
        sc, _ = molecular_surface_scene(n_atoms=int(rng.integers(300, 1200)), width=width, height=height,
                                        ao_samples=0, alpha=1.0, seed=seed)
    d = rng.normal(size=3)
    d[2] = -abs(d[2]) - 0.3
    d /= np.linalg.norm(d)
    light = sc.lights[0]
    sc.lights[0] = replace(light, direction=tuple(float(x) for x in d), intensity=float(rng.uniform(0.5, 1.0)))
    a = float(rng.uniform(0.4, 0.9))
    sc.ambient_color = (a, a, a)
    sc.fog.enabled = False
    return sc


def render_features(dev: Device, bs: BuiltScene, sc: Scene, opt: RenderOptions) -> np.ndarray:
    return bs.render(sc, opt, hires_only=True, want_gi_features=True).gi_features


def make_pairs(dev: Device, seeds: List[int], size: int = 256, target_spp: int = 1024, log=None):
    """(features (n, size, size, 10), targets (n, size, size, 3)): noisy
    features at 1-8 spp, targets = log(1 + E) at target_spp."""
    log = log or _say
    feats, tgts = [], []
    t0 = time.perf_counter()
    for i, s in enumerate(seeds):
        sc = training_scene(s, size, size)
        bs = BuiltScene(dev, sc)
        try:
            spp = int(2 ** np.random.default_rng(s + 7).integers(0, 4))
            f = render_features(dev, bs, sc, gi_options(size, size, spp, seed=1000 + s))
            t = render_features(dev, bs, sc, gi_options(size, size, target_spp, seed=5000 + s))
        finally:
            bs.close()
        feats.append(f)
        tgts.append(t[..., 0:3].copy())
        if (i + 1) % 16 == 0:
            log(f"pairs {i + 1}/{len(seeds)} ({time.perf_counter() - t0:.1f} s)")
    return np.stack(feats), np.stack(tgts)


def train(dev: Device, feats: np.ndarray, tgts: np.ndarray, steps: int = 3000, batch: int = 8, crop: int = 128,
          lr: float = 1e-3, seed: int = 0, levels: int = DEFAULT_LEVELS, channels=DEFAULT_CHANNELS,
          log=None) -> Unet:
    """Adam on random crops (with flips), cosine-decayed learning rate."""
    log = log or _say
    net = Unet(dev, levels, channels, seed=seed)
    rng = np.random.default_rng(seed)
    n, h, w, _ = feats.shape
    t0 = time.perf_counter()
    avg = None
    for it in range(steps):
        idx = rng.integers(0, n, batch)
        ys = rng.integers(0, h - crop + 1, batch)
        xs = rng.integers(0, w - crop + 1, batch)
        fb = np.stack([feats[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)])
        tb = np.stack([tgts[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)])
        if rng.integers(0, 2):
            fb, tb = fb[:, :, ::-1], tb[:, :, ::-1]
            fb = fb.copy()
            fb[..., 6] = -fb[..., 6]  # a mirrored image mirrors the normals' x
        cur = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * it / max(1, steps))))
        loss, _ = net.train_step(fb, tb, cur)
        avg = loss if avg is None else 0.98 * avg + 0.02 * loss
        if (it + 1) % 250 == 0:
            log(f"step {it + 1}/{steps}: loss {avg:.5f} ({time.perf_counter() - t0:.1f} s)")
    return net


def save(path: str, net: Unet) -> None:
    np.savez_compressed(path, params=net.get_params().astype(np.float16), levels=net.levels,
                        channels=np.array(net.channels, np.int32))


def load(dev: Device, path: str) -> Unet:
    z = np.load(path)
    net = Unet(dev, int(z["levels"]), [int(c) for c in z["channels"]], seed=0)
    net.set_params(z["params"].astype(np.float32))
    return net


def time_to_psnr(points: List[Tuple[float, ...]], target: float, col: int = 0) -> Optional[float]:
    """The shortest frame time (ms, column col of (wall ms, dB, device ms))
    among the points whose PSNR reaches target; None when none does."""
    ok = [p[col] for p in points if p[1] >= target]
    return min(ok) if ok else None


def evaluate(dev: Device, net: Optional[Unet], seeds: List[int], size: int = 512, ref_spp: int = 4096,
             plain_spp=(1, 2, 4, 8, 16, 32, 64, 128, 256), den_spp=(1, 2, 4, 8), repeats: int = 3,
             log=None) -> dict:
    """Per held-out scene: frame time and PSNR (sRGB, full frame) against a
    ref_spp render for the plain path tracer, the a-trous denoiser and the
    learned one; the time each first reaches 40 dB."""
    log = log or _say
    out = []
    for s in seeds:
        sc = training_scene(s, size, size)
        bs = BuiltScene(dev, sc)
        try:
            ref = bs.render(sc, gi_options(size, size, ref_spp, seed=9000 + s)).color

            def timed(opt, **kw):
                # wall time of the whole frame call (colour copied to the
                # host) and the device's frame window (trace + shade +
                # gather + denoise + fog + resolve)
                best, dev_ms, f = None, None, None
                for _ in range(repeats):
                    t = time.perf_counter()
                    f = bs.render(sc, opt, color_only=True, **kw)
                    ms = (time.perf_counter() - t) * 1e3
                    best = ms if best is None else min(best, ms)
                    d = (f.render_seconds + f.post_seconds) * 1e3
                    dev_ms = d if dev_ms is None else min(dev_ms, d)
                return best, psnr_srgb(f.color, ref), dev_ms

            plain = [timed(gi_options(size, size, n, seed=100 + s)) for n in plain_spp]
            atrous = [timed(replace(gi_options(size, size, n, seed=100 + s), pt1_denoise=True)) for n in den_spp]
            learned = [timed(gi_options(size, size, n, seed=100 + s), denoise_net=net) for n in den_spp] if net else []
        finally:
            bs.close()
        row = {"seed": s, "plain": plain, "atrous": atrous, "learned": learned,
               "t40_plain": time_to_psnr(plain, 40.0), "t40_atrous": time_to_psnr(atrous, 40.0),
               "t40_learned": time_to_psnr(learned, 40.0) if net else None,
               "t40_plain_device": time_to_psnr(plain, 40.0, 2), "t40_atrous_device": time_to_psnr(atrous, 40.0, 2),
               "t40_learned_device": time_to_psnr(learned, 40.0, 2) if net else None}
        log(f"scene {s}: plain {[(n, round(p[1], 2)) for n, p in zip(plain_spp, plain)]}")
        log(f"  atrous {[(n, round(p[1], 2)) for n, p in zip(den_spp, atrous)]}")
        if net:
            log(f"  learned {[(n, round(p[1], 2)) for n, p in zip(den_spp, learned)]}")
        log(f"  time to 40 dB (wall / device ms): plain {row['t40_plain']} / {row['t40_plain_device']}, "
            f"a-trous {row['t40_atrous']} / {row['t40_atrous_device']}, "
            f"learned {row['t40_learned']} / {row['t40_learned_device']}")
        out.append(row)
    return {"size": size, "ref_spp": ref_spp, "scenes": out}


def bench_infer(dev: Device, weights: str, size: int, calls: int) -> float:
    """Wall ms per U-Net inference on one size x size image of seeded random
    features (host copies included); the kernels' own times come from a
    kernel trace of this run."""
    net = load(dev, weights)
    rng = np.random.default_rng(0)
    f = rng.uniform(0.0, 1.0, size=(1, size, size, 10)).astype(np.float32)
    net.infer(f)
    t = time.perf_counter()
    for _ in range(calls):
        net.infer(f)
    ms = (time.perf_counter() - t) * 1e3 / max(1, calls)
    _say(f"unet inference {size}x{size}: {ms:.3f} ms per call over {calls} calls")
    net.close()
    return ms


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--scenes", type=int, default=96)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default="")
    ap.add_argument("--evaluate", default="", help="weights to evaluate on held-out scenes")
    ap.add_argument("--eval-scenes", type=int, default=3)
    ap.add_argument("--eval-size", type=int, default=512)
    ap.add_argument("--bench-infer", type=int, default=0, metavar="SIZE",
                    help="time inference of the shipped weights on one SIZE x SIZE image")
    ap.add_argument("--calls", type=int, default=11)
    args = ap.parse_args()
    dev = Device(0)
    try:
        if args.bench_infer:
            import os
            bench_infer(dev, os.path.join(os.path.dirname(__file__), "weights", "gi_unet.npz"),
                        args.bench_infer, args.calls)
            return
        if args.evaluate:
            net = load(dev, args.evaluate)
            evaluate(dev, net, [100000 + i for i in range(args.eval_scenes)], size=args.eval_size)
            return
        feats, tgts = make_pairs(dev, list(range(args.scenes)), size=args.size, target_spp=args.target_spp)
        net = train(dev, feats, tgts, steps=args.steps)
        if args.out:
            save(args.out, net)
            _say(f"saved {args.out} ({net.n_params} parameters)")
    finally:
        dev.close()


if __name__ == "__main__":
    main()
