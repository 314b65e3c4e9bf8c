"""Batched render serving over ubg_render_batch.

The reference serves one render per call (src/umbreon/umbreon.cpp:322-341,
renderAsync: one worker thread per request). Here a server keeps built scenes
resident in HBM, admits the requests that have arrived into batches by an HBM
budget, and renders each batch through one set of wavefront launches
(``umbreon_amd.render.render_batch``). Per-request images equal the
single-request renders whatever the batch composition or tuning.

* ``request_stream``: the seeded workload -- Poisson arrivals, log-uniform
  atom counts, mixed representations (CPK spheres, ball-and-stick, molecular
  surface), image sizes from {512x512, 1024x768, 1920x1080}, AO 16, a camera
  orbit angle per request.
* ``MoleculePool``: the synthetic molecules behind the stream (generated on
  the host before serving, like a client's uploaded scene).
* ``Server``: scene cache (GPU BVH build on first use, LRU eviction within the
  HBM budget), admission in arrival order while the batch's frame scratch and
  scenes fit the budget, latency bookkeeping.
* ``autotune``: picks the BVH8 leaf size per representation and the AO
  waves per SIMD and batch size on calibration requests.
"""
from __future__ import annotations

import copy
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import capi
from .render import BuiltScene, Device, render_batch
from .scene import Camera, RenderOptions, Scene

SIZES = ((512, 512), (1024, 768), (1920, 1080))
REPS = ("spheres", "ballstick", "surface")


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    atoms: int              # pool molecule size
    rep: str
    width: int
    height: int
    angle: float            # camera orbit about the molecule's vertical axis, radians
    seed: int = 0


def pool_sizes(n: int = 8, lo: int = 500, hi: int = 200_000) -> List[int]:
    """Log-spaced molecule sizes covering [lo, hi]."""
    return [int(round(lo * (hi / lo) ** (k / (n - 1)))) for k in range(n)]


def request_stream(n: int = 256, rate: float = 40.0, seed: int = 0, surface_max_atoms: int = 40_000,
                   sizes: Optional[Sequence[int]] = None) -> List[Request]:
    """`n` requests with exponential inter-arrival times (mean 1 / rate s),
    atom counts log-uniform in [500, 200000] snapped to the pool's nearest
    size (surfaces to the sizes <= surface_max_atoms), representation and
    image size uniform."""
    rng = np.random.default_rng(seed)
    pool = list(sizes) if sizes is not None else pool_sizes()
    surf = [s for s in pool if s <= surface_max_atoms]
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        atoms = float(math.exp(rng.uniform(math.log(500.0), math.log(200_000.0))))
        rep = REPS[int(rng.integers(0, len(REPS)))]
        cand = surf if rep == "surface" else pool
        snap = min(cand, key=lambda s: abs(math.log(s) - math.log(atoms)))
        w, h = SIZES[int(rng.integers(0, len(SIZES)))]
        out.append(Request(i, t, snap, rep, w, h, float(rng.uniform(0.0, 2.0 * math.pi))))
    return out


class MoleculePool:
    """Host-side scenes of the stream's molecules, generated once."""

    def __init__(self):
        self.scenes: Dict[Tuple[int, str], Scene] = {}

    def get(self, atoms: int, rep: str) -> Scene:
        key = (atoms, rep)
        if key not in self.scenes:
            from .scenes import ball_and_stick_scene, molecular_surface_scene
            if rep == "spheres":
                sc, _ = _cpk_scene(atoms)
            elif rep == "ballstick":
                sc, _ = ball_and_stick_scene(n_atoms=atoms, width=1024, height=768, supersample=1)
            else:
                sc, _ = molecular_surface_scene(n_atoms=atoms, width=1024, height=768)
            self.scenes[key] = sc
        return self.scenes[key]


def _cpk_scene(atoms: int):
    from .scenes import spheres_scene
    sc, opt = spheres_scene(n_atoms=atoms, width=1024, height=768)
    return sc, opt


def request_view(sc: Scene, req: Request) -> Tuple[Scene, RenderOptions]:
    """The request's camera (the scene's framing camera orbited by the
    request's angle about the vertical axis through the framed centre) and
    options: AO 16 (8 A), hard shadows, 1 spp, the scene's transparency."""
    v = copy.copy(sc)
    c = sc.camera
    pos = np.asarray(c.position, np.float64)
    d = np.asarray(c.direction, np.float64)
    dist = float(np.linalg.norm(d)) or 1.0
    # centre: the point the framing camera looks at, at its framing distance
    # (frame_camera puts the camera on +z of the centre)
    centre = pos.copy()
    centre[2] -= _framing_distance(sc)
    ca, sa = math.cos(req.angle), math.sin(req.angle)
    rel = pos - centre
    rp = np.array([ca * rel[0] + sa * rel[2], rel[1], -sa * rel[0] + ca * rel[2]])
    npos = centre + rp
    nd = (centre - npos) / max(1e-9, float(np.linalg.norm(centre - npos))) * dist
    v.camera = Camera(position=tuple(float(x) for x in npos), direction=tuple(float(x) for x in nd),
                      up=c.up, fovy=c.fovy, orthographic=c.orthographic, height=c.height)
    opt = RenderOptions(width=req.width, height=req.height, supersample=1, shadows=True, ao_samples=16,
                        ao_distance=8.0)
    return v, opt


def _framing_distance(sc: Scene) -> float:
    """Distance of the framing camera from the centre it frames: the camera
    sits on +z of the bounding-box centre of the geometry."""
    pts = []
    if len(sc.sphere_center):
        pts.append(sc.sphere_center)
    if sc.positions is not None and len(sc.positions):
        pts.append(sc.positions)
    p = np.concatenate(pts) if pts else np.zeros((1, 3), np.float32)
    lo, hi = p.min(axis=0), p.max(axis=0)
    centre_z = 0.5 * (float(lo[2]) + float(hi[2]))
    return float(sc.camera.position[2]) - centre_z


def frame_bytes(width: int, height: int) -> int:
    """Frame scratch of one request in the batch pool (queues, walk state,
    outputs; ~150 bytes per pixel at 1 spp)."""
    from . import capi as _c  # noqa: F401
    px = width * height
    return px * 68 + px * 56 + px * 16 + 64 * 1024


@dataclass
class Tuning:
    ao_waves_per_simd: int = 5
    max_batch: int = 64
    leaf_max: Dict[str, int] = field(default_factory=lambda: {r: 3 for r in REPS})


class Server:
    """Scene cache + HBM-budget admission + batched rendering."""

    def __init__(self, device: Device, pool: MoleculePool, hbm_budget: int = 96 << 30,
                 tuning: Optional[Tuning] = None):
        self.device = device
        self.pool = pool
        self.budget = int(hbm_budget)
        self.tuning = tuning or Tuning()
        self.cache: Dict[Tuple[int, str, int], BuiltScene] = {}
        self.last_use: Dict[Tuple[int, str, int], float] = {}
        self.build_seconds = 0.0

    def close(self) -> None:
        for b in self.cache.values():
            b.close()
        self.cache.clear()

    def scene_bytes(self) -> int:
        return sum(int(b.info.get("device_bytes", 0)) for b in self.cache.values())

    def built(self, atoms: int, rep: str, pinned: set) -> BuiltScene:
        leaf = int(self.tuning.leaf_max.get(rep, 3))
        key = (atoms, rep, leaf)
        if key not in self.cache:
            sc = self.pool.get(atoms, rep)
            t0 = time.perf_counter()
            self.cache[key] = BuiltScene(self.device, sc, capi.BVH8_GPU_BUILD, leaf_max=leaf)
            self.build_seconds += time.perf_counter() - t0
            # evict least recently used scenes beyond the budget (never one the batch uses)
            while self.scene_bytes() > self.budget // 2:
                victims = [k for k in self.cache if k not in pinned and k != key]
                if not victims:
                    break
                v = min(victims, key=lambda k: self.last_use.get(k, 0.0))
                self.cache.pop(v).close()
                self.last_use.pop(v, None)
        self.last_use[key] = time.perf_counter()
        return self.cache[key]

    def admit(self, pending: List[Request]) -> List[Request]:
        """The longest arrival-order prefix of `pending` whose frame scratch
        fits the budget left by the resident scenes, at most max_batch."""
        room = self.budget - self.scene_bytes()
        batch, used = [], 0
        for r in pending:
            fb = frame_bytes(r.width, r.height)
            if batch and (used + fb > room or len(batch) >= self.tuning.max_batch):
                break
            batch.append(r)
            used += fb
        return batch

    def render(self, batch: List[Request], want_images: bool = True):
        pinned = {(r.atoms, r.rep, int(self.tuning.leaf_max.get(r.rep, 3))) for r in batch}
        built, scenes, opts = [], [], []
        for r in batch:
            b = self.built(r.atoms, r.rep, pinned)
            sc, opt = request_view(self.pool.get(r.atoms, r.rep), r)
            built.append(b)
            scenes.append(sc)
            opts.append(opt)
        return render_batch(self.device, built, scenes, opts, color_only=True,
                            ao_waves_per_simd=self.tuning.ao_waves_per_simd)

    def run(self, stream: List[Request], log=None) -> dict:
        """Serves the stream in real time: requests become visible at their
        arrival times; each batch takes every visible request the budget
        admits. Returns throughput and latency figures."""
        pending: List[Request] = []
        i = 0
        done: Dict[int, float] = {}
        rays = 0
        batches = 0
        t0 = time.perf_counter()
        while len(done) < len(stream):
            now = time.perf_counter() - t0
            while i < len(stream) and stream[i].arrival <= now:
                pending.append(stream[i])
                i += 1
            if not pending:
                time.sleep(max(0.0, stream[i].arrival - now))
                continue
            batch = self.admit(pending)
            pending = pending[len(batch):]
            frames = self.render(batch)
            t = time.perf_counter() - t0
            for r, f in zip(batch, frames):
                done[r.rid] = t
                rays += f.total_rays
            batches += 1
            if log:
                log(f"batch {batches}: {len(batch)} requests, t={t:.3f}s, {len(pending)} waiting")
        lat = np.array([done[r.rid] - r.arrival for r in stream])
        first = min(r.arrival for r in stream)
        last = max(done.values())
        return {"requests": len(stream), "batches": batches, "images_per_s": len(stream) / (last - first),
                "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_latency_ms": float(lat.mean() * 1e3), "makespan_s": last - first,
                "mrays_per_s": rays / (last - first) / 1e6, "scene_build_s": self.build_seconds}


def offline_throughput(server: Server, reqs: List[Request]) -> float:
    """Images per second of rendering `reqs` back to back in admitted
    batches (scenes already built): the autotuner's objective."""
    todo = list(reqs)
    for r in todo:
        server.built(r.atoms, r.rep, set())
    t0 = time.perf_counter()
    while todo:
        b = server.admit(todo)
        todo = todo[len(b):]
        server.render(b)
    return len(reqs) / (time.perf_counter() - t0)


def autotune(device: Device, pool: MoleculePool, calib: List[Request], budget: int, log=None) -> Tuning:
    """Tuning by measurement: per representation the BVH8 leaf size (1..3)
    that renders its largest calibration molecule fastest, then the AO waves
    per SIMD and the batch size that serve the calibration requests fastest.
    Images do not depend on any of these."""
    tun = Tuning()
    for rep in REPS:
        rs = [r for r in calib if r.rep == rep]
        if not rs:
            continue
        big = max(rs, key=lambda r: r.atoms)
        best = None
        for leaf in (1, 2, 3):
            b = BuiltScene(device, pool.get(big.atoms, rep), capi.BVH8_GPU_BUILD, leaf_max=leaf)
            try:
                sc, opt = request_view(pool.get(big.atoms, rep), big)
                render_batch(device, [b], [sc], [opt], color_only=True)
                t0 = time.perf_counter()
                for _ in range(3):
                    render_batch(device, [b], [sc], [opt], color_only=True)
                dt = time.perf_counter() - t0
            finally:
                b.close()
            if log:
                log(f"autotune {rep} ({big.atoms} atoms): leaf {leaf}: {dt / 3 * 1e3:.2f} ms/frame")
            if best is None or dt < best[0]:
                best = (dt, leaf)
        tun.leaf_max[rep] = best[1]
    best = None
    for waves in (3, 5):
        for mb in (16, 64):
            t = Tuning(ao_waves_per_simd=waves, max_batch=mb, leaf_max=dict(tun.leaf_max))
            srv = Server(device, pool, budget, t)
            try:
                offline_throughput(srv, calib)  # warm-up (builds, first launches)
                ips = offline_throughput(srv, calib)
            finally:
                srv.close()
            if log:
                log(f"autotune waves/SIMD {waves}, batch {mb}: {ips:.1f} images/s")
            if best is None or ips > best[0]:
                best = (ips, waves, mb)
    tun.ao_waves_per_simd, tun.max_batch = best[1], best[2]
    return tun
