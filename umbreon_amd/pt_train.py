"""The learned denoiser of the path-traced GI workload (ubg_pathtrace):
training pairs the renderer generates itself, training on the device, and
the time-to-40-dB evaluation on the full workload.

Workload (scenes.pathtraced_surface_scene): the Gaussian isosurface of a
40,000-atom synthetic protein (4.03 M triangles) as a GGX dielectric over its
CPK spheres (a GGX conductor or dielectric per element, roughness uniform in
[0.05, 0.6]), 64 spherical area lights by RIS next-event estimation with
ReSTIR spatial reuse, 4 path vertices, 1920x1080, 1-8 samples per pixel.

Training pairs: seeded molecules (3,000-30,000 atoms, the same generator
with their own materials and lights) seen through 512x512 windows of the
full-frame camera rule -- the 1920x1080, 40-degree framing of the molecule,
from a random side, narrowed to the window's share of that frame and aimed
at a random point of the surface -- so a pixel covers about what it covers
in the full frame. Noisy inputs are the path tracer's 12 denoiser features
(encoded colour, guides, and the noise descriptors 1 / sqrt(spp) and the
colour's standard error) at 1, 2, 4 and 8 samples per pixel; the target is
the encoded colour
log(1 + 32 c) / 4 at 4096 samples. The network (5 levels, 32-128 channels,
bf16 MFMA convolutions with fp32 accumulation) predicts the encoded colour
from the features; the L1 loss on that encoding weighs dark colours about as
the sRGB comparison does.

    python -m umbreon_amd.pt_train --gen-seconds 600 --train-seconds 420 --out w.npz [--init w0.npz]
    python -m umbreon_amd.pt_train --evaluate umbreon_amd/weights/pt_unet.npz
    python -m umbreon_amd.pt_train --bench-frame 1 --weights umbreon_amd/weights/pt_unet.npz --calls 5
"""
from __future__ import annotations

import argparse
import math
import multiprocessing
import time
from dataclasses import replace
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from . import capi
from .denoise import Unet
from .denoise_train import load, save, time_to_psnr
from .image_io import psnr_srgb
from .pathtrace import pathtrace
from .render import BuiltScene, Device
from .scene import Camera
from .scenes import pathtraced_surface_scene

LEVELS = 5
CHANNELS = (32, 64, 96, 128, 128)
FULL_H = 1080          # rows of the full frame whose pixel footprint the windows copy
FULL_FOVY = 40.0       # its vertical field of view (degrees)
WINDOW = 512
NOISY_SPP = (1, 2, 4, 8)
TARGET_SPP = 4096
WEIGHTS = Path(__file__).resolve().parent / "weights" / "pt_unet.npz"


def encode_colour(c: np.ndarray) -> np.ndarray:
    """The path tracer's colour encoding of the denoiser's features and
    prediction, u = log(1 + 32 c) / 4 (sRGB-like weighting of dark colours)."""
    return (0.25 * np.log1p(32.0 * np.maximum(np.asarray(c, np.float64), 0.0))).astype(np.float32)


def _say(msg: str) -> None:
    print(msg, flush=True)


def _rotation(rng: np.random.Generator) -> np.ndarray:
    """A uniformly random rotation (unit quaternion from four normals)."""
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def window_scene(seed: int, size: int = WINDOW):
    """One training molecule and a size x size window camera on it; returns
    (Scene, PtOptions). The camera stands where the full-frame framing would
    put it (bounding radius / sin(20 deg) * 1.02 from the centre, the rule of
    scenes.frame_camera) on a random side, with the field of view of size
    rows of a 1080-row, 40-degree frame, aimed at a random surface vertex."""
    rng = np.random.default_rng([seed, 0x5EED])
    n_atoms = int(round(math.exp(rng.uniform(math.log(3000.0), math.log(30000.0)))))
    sc, opt = pathtraced_surface_scene(n_atoms=n_atoms, width=size, height=size, seed=seed)
    v = np.asarray(sc.positions, np.float64)
    c = 0.5 * (v.min(axis=0) + v.max(axis=0))
    r = float(np.max(np.linalg.norm(v - c, axis=1)))
    dist = r / math.sin(math.radians(FULL_FOVY) * 0.5) * 1.02
    R = _rotation(rng)
    pos = c + R @ np.array([0.0, 0.0, dist])
    d = v[int(rng.integers(0, len(v)))] - pos
    d /= np.linalg.norm(d)
    up = R @ np.array([0.0, 1.0, 0.0])
    fovy = 2.0 * math.degrees(math.atan(math.tan(math.radians(FULL_FOVY) * 0.5) * size / FULL_H))
    sc.camera = Camera(position=tuple(float(a) for a in pos), direction=tuple(float(a) for a in d),
                       up=tuple(float(a) for a in up), fovy=fovy)
    return sc, opt


def scene_pool(workers: int = 4):
    """Host processes for window_scene (the generator is Python-bound). Make
    it before the first device call: the workers are fresh interpreters
    started from a process that has not touched the GPU."""
    return multiprocessing.get_context("spawn").Pool(workers)


def make_pairs(dev: Device, seeds: Sequence[int], size: int = WINDOW, target_spp: int = TARGET_SPP,
               noisy_spp: Sequence[int] = NOISY_SPP, seconds: Optional[float] = None, pool=None, log=None):
    """Render training pairs until the seeds or the time budget run out.

    Returns (features (n * len(noisy_spp), size, size, 12) float16, owner
    (n * len(noisy_spp),) index of each input's target, targets (n, size,
    size, 3) float32 encoded colour). With a scene_pool the next molecules
    are generated on the host while the device renders."""
    log = log or _say
    feats, owner, tgts = [], [], []
    t0 = time.perf_counter()
    seeds = list(seeds)
    ahead = 6
    pending = [pool.apply_async(window_scene, (s, size)) for s in seeds[:ahead]] if pool is not None else []
    for i, s in enumerate(seeds):
        if pool is not None:
            sc, opt = pending.pop(0).get()
            if i + ahead < len(seeds):
                pending.append(pool.apply_async(window_scene, (seeds[i + ahead], size)))
        else:
            sc, opt = window_scene(s, size)
        bs = BuiltScene(dev, sc, capi.BVH8_GPU_BUILD)
        try:
            t = pathtrace(bs, sc, replace(opt, spp=target_spp, seed=(7919 * s + 1) & 0xFFFFFFFF))
            tgts.append(encode_colour(t.color[..., :3]))
            for k, n in enumerate(noisy_spp):
                f = pathtrace(bs, sc, replace(opt, spp=n, seed=(7919 * s + 11 + k) & 0xFFFFFFFF),
                              want_features=True)
                feats.append(f.features.astype(np.float16))
                owner.append(len(tgts) - 1)
        finally:
            bs.close()
        el = time.perf_counter() - t0
        log(f"pair {i + 1}: seed {s}, {len(sc.sphere_center)} atoms, {len(sc.index)} triangles, "
            f"{t.rays / 1e6:.0f} M target rays ({t.trace_seconds:.2f} s), {el:.0f} s")
        if seconds is not None and el >= seconds:
            break
    return np.stack(feats), np.asarray(owner, np.int64), np.stack(tgts)


def _dihedral(a: np.ndarray, k: int, normals: bool = False) -> np.ndarray:
    """One of the 8 flips / quarter turns of (batch, h, w, c) crops. With
    normals, channels 6-7 are the camera-space normal's (right, up)
    components and turn with the image: columns run along +right and rows
    along -up, so a transpose maps (x, y) to (-y, -x), a column flip negates
    x and a row flip negates y."""
    if k & 4:
        a = a.transpose(0, 2, 1, 3)
        if normals:
            a = a.copy()
            a[..., [6, 7]] = -a[..., [7, 6]]
    if k & 1:
        a = a[:, :, ::-1]
        if normals:
            a = a.copy()
            a[..., 6] = -a[..., 6]
    if k & 2:
        a = a[:, ::-1]
        if normals:
            a = a.copy()
            a[..., 7] = -a[..., 7]
    return a


def train(dev: Device, feats: np.ndarray, owner: np.ndarray, tgts: np.ndarray, seconds: float = 300.0,
          batch: int = 8, crop: int = 128, lr: float = 1e-3, net: Optional[Unet] = None, seed: int = 0,
          log=None) -> Unet:
    """Adam on random crops (each under one of the 8 flips / quarter turns,
    the camera-space normals turned with it) for about `seconds` of wall
    time, cosine-decayed learning rate over the step count the first steps'
    pace allows."""
    log = log or _say
    if net is None:
        net = Unet(dev, LEVELS, CHANNELS, seed=seed)
    rng = np.random.default_rng(seed + 12345)
    n, h, w, _ = feats.shape

    def batch_at():
        idx = rng.integers(0, n, batch)
        ys = rng.integers(0, h - crop + 1, batch)
        xs = rng.integers(0, w - crop + 1, batch)
        fb = np.stack([feats[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)]).astype(np.float32)
        tb = np.stack([tgts[owner[i], y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)])
        k = int(rng.integers(0, 8))
        return np.ascontiguousarray(_dihedral(fb, k, normals=True)), np.ascontiguousarray(_dihedral(tb, k))

    t0 = time.perf_counter()
    steps, it, avg = None, 0, None
    while True:
        total = steps if steps is not None else 10 ** 9
        cur = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * min(1.0, it / total))))
        fb, tb = batch_at()
        loss, _ = net.train_step(fb, tb, cur)
        avg = loss if avg is None else 0.98 * avg + 0.02 * loss
        it += 1
        if it == 30:
            pace = (time.perf_counter() - t0) / it
            steps = max(it + 1, int(seconds / pace))
            log(f"training: {pace * 1e3:.1f} ms per step of {batch} crops of {crop}^2, {steps} steps planned")
        if it % 500 == 0:
            log(f"step {it}/{steps}: loss {avg:.5f} ({time.perf_counter() - t0:.0f} s)")
        if steps is not None and it >= steps:
            break
    log(f"trained {it} steps in {time.perf_counter() - t0:.0f} s, final loss {avg:.5f}")
    return net


def holdout_psnr(dev: Device, net: Unet, seeds: Sequence[int], size: int = WINDOW, ref_spp: int = 1024,
                 log=None) -> List[dict]:
    """Windows of held-out molecules: sRGB PSNR of the plain and the denoised
    image at 1-8 samples against a ref_spp render."""
    log = log or _say
    out = []
    for s in seeds:
        sc, opt = window_scene(s, size)
        bs = BuiltScene(dev, sc, capi.BVH8_GPU_BUILD)
        try:
            ref = pathtrace(bs, sc, replace(opt, spp=ref_spp, seed=99)).color
            row = {"seed": s, "plain": [], "learned": []}
            for n in NOISY_SPP:
                o = replace(opt, spp=n, seed=5)
                row["plain"].append(round(psnr_srgb(pathtrace(bs, sc, o).color, ref), 2))
                row["learned"].append(round(psnr_srgb(pathtrace(bs, sc, o, denoise_net=net).color, ref), 2))
        finally:
            bs.close()
        log(f"held-out window {s}: plain {row['plain']} dB, learned {row['learned']} dB at {list(NOISY_SPP)} spp")
        out.append(row)
    return out


# samples per pixel of the evaluation's frames: steps of 2^(1/2) (rounded);
# each crossing is then refined between the last step below it and the first
# at or above it (refine_crossings), to within 2^(1/8)
SPP_STEPS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048)
THRESHOLDS_DB = (30.0, 35.0, 40.0)


def refine_crossings(points: dict, measure, thresholds: Sequence[float] = THRESHOLDS_DB, log=None) -> None:
    """points: {spp: (wall ms, dB, ...)} of a sweep over SPP_STEPS. For each
    threshold that some step reaches, except at the first step, measure
    (spp -> point) the three sample counts 2^(1/8) apart between the last
    step below it and the first step at or above it, in increasing order,
    until one reaches it; the points join `points`. The same rule serves
    both methods, so neither's crossing is rounded up further than the
    other's."""
    log = log or _say
    for db in thresholds:
        steps = sorted(points)
        ok = [n for n in steps if points[n][1] >= db]
        if not ok or ok[0] == steps[0]:
            continue
        hi = ok[0]
        lo = max(n for n in steps if n < hi)
        cand = sorted({int(round(lo * (hi / lo) ** (i / 4))) for i in (1, 2, 3)} - {lo, hi})
        for n in cand:
            if n not in points:
                points[n] = measure(n)
                log(f"  refine {db:g} dB: {n} spp: {points[n][1]:.2f} dB, {points[n][0]:.1f} ms wall")
            if points[n][1] >= db:
                break


def evaluate_full(dev: Device, net: Optional[Unet], ref_spp: int = TARGET_SPP, max_plain_spp: int = 2048,
                  den_spp: Sequence[int] = SPP_STEPS, repeats: int = 3, log=None) -> dict:
    """Time to 40 dB on the full workload (1920x1080, seed 0, a molecule the
    training never saw): PSNR (sRGB, whole frame) against a ref_spp render;
    the plain path tracer at 1, 2, 3, 4, 6, 8, ... samples until it reaches
    40 dB (at most max_plain_spp), the learned denoiser at den_spp until it
    does, each crossing of 30, 35 and 40 dB then refined to within 2^(1/8)
    samples (refine_crossings); the time each first reaches 30, 35 and
    40 dB. The denoiser runs with its default ensemble (8 network passes
    from 32 spp on). Each point is
    (wall ms of the frame call with the colour copied to the host, dB, device
    ms = sample passes + resolve + network)."""
    log = log or _say
    t0 = time.perf_counter()
    sc, opt = pathtraced_surface_scene()
    gen_s = time.perf_counter() - t0
    bs = BuiltScene(dev, sc, capi.BVH8_GPU_BUILD)
    try:
        log(f"full workload: {len(sc.index)} triangles + {len(sc.sphere_center)} spheres, "
            f"{len(opt.lights)} lights, generated in {gen_s:.1f} s, GPU BVH build "
            f"{bs.info['build_seconds'] * 1e3:.0f} ms")
        t = time.perf_counter()
        ref = pathtrace(bs, sc, replace(opt, spp=ref_spp, seed=777)).color
        log(f"reference: {ref_spp} spp in {time.perf_counter() - t:.1f} s")

        def timed(o, **kw):
            best, dev_ms, den_ms, f = None, None, None, None
            for _ in range(repeats if o.spp <= 64 else 1):
                t = time.perf_counter()
                f = pathtrace(bs, sc, o, **kw)
                ms = (time.perf_counter() - t) * 1e3
                d = (f.trace_seconds + f.denoise_seconds) * 1e3
                best = ms if best is None else min(best, ms)
                dev_ms = d if dev_ms is None else min(dev_ms, d)
                den_ms = f.denoise_seconds * 1e3 if den_ms is None else min(den_ms, f.denoise_seconds * 1e3)
            return (best, psnr_srgb(f.color, ref), dev_ms, den_ms)

        plain_pts = {}
        for n in SPP_STEPS:
            if n > max_plain_spp:
                break
            plain_pts[n] = timed(replace(opt, spp=n, seed=11))
            log(f"plain {n} spp: {plain_pts[n][1]:.2f} dB, {plain_pts[n][0]:.1f} ms wall, "
                f"{plain_pts[n][2]:.1f} ms device")
            if plain_pts[n][1] >= 40.0:
                break
        refine_crossings(plain_pts, lambda n: timed(replace(opt, spp=n, seed=11)), log=log)
        learned_pts = {}
        if net is not None:
            for n in den_spp:
                if n > max(plain_pts):
                    break  # no faster than the plain path tracer any more
                learned_pts[n] = timed(replace(opt, spp=n, seed=11), denoise_net=net)
                p = learned_pts[n]
                log(f"learned {n} spp: {p[1]:.2f} dB, {p[0]:.1f} ms wall, {p[2]:.1f} ms device "
                    f"({p[3]:.2f} ms resolve + network)")
                if p[1] >= 40.0:
                    break
            refine_crossings(learned_pts, lambda n: timed(replace(opt, spp=n, seed=11), denoise_net=net), log=log)
        plain_spp, plain = sorted(plain_pts), [plain_pts[n] for n in sorted(plain_pts)]
        den_done, learned = sorted(learned_pts), [learned_pts[n] for n in sorted(learned_pts)]
    finally:
        bs.close()
    res = {"workload": (f"path-traced GI: molecular surface ({len(sc.index)} triangles) as a GGX dielectric over "
                        f"{len(sc.sphere_center)} CPK spheres (GGX conductors / dielectrics, roughness "
                        f"U[0.05, 0.6]), {len(opt.lights)} spherical area lights (RIS + ReSTIR spatial), "
                        f"{opt.bounces} bounces, 1920x1080; PSNR (sRGB, whole frame) vs {ref_spp} spp"),
           "ref_spp": ref_spp,
           "points": "spp: [wall ms, dB, device ms, of which resolve + network ms]",
           "plain": {str(n): [round(float(v), 2) for v in p] for n, p in zip(plain_spp, plain)},
           "learned": {str(n): [round(float(v), 2) for v in p] for n, p in zip(den_done, learned)},
           "t40_plain_ms": time_to_psnr(plain, 40.0), "t40_learned_ms": time_to_psnr(learned, 40.0),
           "t40_plain_device_ms": time_to_psnr(plain, 40.0, 2),
           "t40_learned_device_ms": time_to_psnr(learned, 40.0, 2)}
    # time (wall ms) each first reaches a PSNR, and the ratio
    res["time_to_db"] = {}
    for db in THRESHOLDS_DB:
        a, b = time_to_psnr(plain, db), time_to_psnr(learned, db)
        res["time_to_db"][f"{db:g}"] = {"plain_ms": a, "learned_ms": b,
                                       "speedup": None if (a is None or b is None) else round(a / b, 2)}
    for a, b, k in (("t40_plain_ms", "t40_learned_ms", "speedup"),
                    ("t40_plain_device_ms", "t40_learned_device_ms", "speedup_device")):
        res[k] = None if (res[a] is None or res[b] is None) else round(res[a] / res[b], 2)
    log(f"time to 40 dB: {res}")
    return res


def bench_frame(dev: Device, weights: Optional[str], spp: int, calls: int, log=None) -> dict:
    """Wall and device ms per full-workload frame (1920x1080) at spp samples
    per pixel, denoised by the network at `weights` (None: plain), after one
    warm-up frame; for kernel traces of the path tracer and the network."""
    log = log or _say
    sc, opt = pathtraced_surface_scene()
    bs = BuiltScene(dev, sc, capi.BVH8_GPU_BUILD)
    net = load(dev, weights) if weights else None
    try:
        o = replace(opt, spp=spp, seed=11)
        pathtrace(bs, sc, o, denoise_net=net)
        walls, traces, dens, rays = [], [], [], 0
        for _ in range(calls):
            t = time.perf_counter()
            f = pathtrace(bs, sc, o, denoise_net=net)
            walls.append((time.perf_counter() - t) * 1e3)
            traces.append(f.trace_seconds * 1e3)
            dens.append(f.denoise_seconds * 1e3)
            rays = f.rays
    finally:
        if net is not None:
            net.close()
        bs.close()
    res = {"spp": spp, "denoised": net is not None, "wall_ms": float(np.median(walls)),
           "trace_ms": float(np.median(traces)), "resolve_network_ms": float(np.median(dens)), "rays": int(rays),
           "mrays_per_s": rays / (float(np.median(traces)) * 1e3)}
    log(f"frame: {res}")
    return res


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--scenes", type=int, default=10000)
    ap.add_argument("--gen-seconds", type=float, default=600.0)
    ap.add_argument("--train-seconds", type=float, default=300.0)
    ap.add_argument("--target-spp", type=int, default=TARGET_SPP)
    ap.add_argument("--noisy-spp", default=",".join(str(n) for n in NOISY_SPP),
                    help="comma-separated samples per pixel of the noisy inputs of every molecule")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", default="", help="weights to continue training from")
    ap.add_argument("--loss", choices=["l1", "l2"], default="l1",
                    help="training loss on the encoded colour: mean absolute (l1) or squared (l2) error")
    ap.add_argument("--out", default="")
    ap.add_argument("--holdout", type=int, default=2, help="held-out windows to report after training")
    ap.add_argument("--evaluate", default="", help="weights to evaluate on the full workload")
    ap.add_argument("--evaluate-after", action="store_true", help="evaluate the trained net on the full workload")
    ap.add_argument("--workers", type=int, default=4, help="host processes generating molecules")
    ap.add_argument("--bench-frame", type=int, default=0, metavar="SPP",
                    help="time full-workload frames at SPP samples per pixel (denoised with --weights)")
    ap.add_argument("--weights", default="", help="network for --bench-frame")
    ap.add_argument("--calls", type=int, default=5)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="with --bench-frame: one device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.bench_frame:
        dev = Device(0)
        try:
            if args.calib_copy_mib > 0:
                from .counters import calib_copy
                calib_copy(args.calib_copy_mib, log=_say)
            bench_frame(dev, args.weights or None, args.bench_frame, args.calls)
        finally:
            dev.close()
        return
    pool = None if args.evaluate else scene_pool(args.workers)
    dev = Device(0)
    try:
        if args.evaluate:
            net = load(dev, args.evaluate)
            evaluate_full(dev, net)
            return
        seeds = range(args.first_seed, args.first_seed + args.scenes)
        noisy = tuple(int(v) for v in args.noisy_spp.split(","))
        feats, owner, tgts = make_pairs(dev, seeds, target_spp=args.target_spp, noisy_spp=noisy,
                                        seconds=args.gen_seconds, pool=pool)
        pool.terminate()
        pool = None
        _say(f"{len(tgts)} molecules, {len(feats)} noisy inputs")
        net = load(dev, args.init) if args.init else Unet(dev, LEVELS, CHANNELS, seed=args.first_seed)
        net.set_loss(args.loss)
        net = train(dev, feats, owner, tgts, seconds=args.train_seconds, batch=args.batch, crop=args.crop,
                    lr=args.lr, net=net, seed=args.first_seed)
        if args.out:
            save(args.out, net)
            _say(f"saved {args.out} ({net.n_params} parameters)")
        if args.holdout:
            holdout_psnr(dev, net, [900000 + i for i in range(args.holdout)])
        if args.evaluate_after:
            evaluate_full(dev, net)
    finally:
        dev.close()
        if pool is not None:
            pool.terminate()


if __name__ == "__main__":
    main()
