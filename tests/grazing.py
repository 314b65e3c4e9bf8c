"""Grazing rays at large coordinates, shared by the CPU and GPU tests of the
FP64 root polish: unit-scale spheres, capped cones and chained round curves
placed ~1e4 from the origin, and rays that pass the silhouette at relative
distances 1e-6 .. 3e-3 inside (true hits) or outside (true misses).

The truth is the FP64 oracle (everything in double from the same float
inputs). `fp32_plain_sphere_hits` is the classic shifted-frame FP32 sphere
test without the polish, used only to show the failure the polish fixes.
"""
from __future__ import annotations

import numpy as np

from umbreon_amd import capi
from umbreon_amd.scene import Scene

OFFSET = np.array([1.0e4, -2.0e4, 1.5e4])


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _perp(v, rng):
    a = rng.normal(size=v.shape)
    a -= (a * v).sum(-1, keepdims=True) * v
    return _unit(a)


def rays(orgs, dirs):
    r = np.zeros(len(orgs), dtype=capi.ray_dtype())
    r["org"] = orgs
    r["dir"] = dirs
    r["tnear"] = 0.0
    r["tfar"] = np.inf
    return r


def sphere_case(rng, n_spheres=40, rays_per=64):
    """Spheres r in [0.5, 2] at ~1e4; each gets rays from ~40 units away
    aimed past its centre at b = r (1 - s delta), s = +-1, delta log-uniform.
    Returns (scene, rays, the index of the sphere each ray grazes)."""
    c = (OFFSET + rng.uniform(-20, 20, (n_spheres, 3))).astype(np.float32)
    r = rng.uniform(0.5, 2.0, n_spheres).astype(np.float32)
    sc = Scene()
    sc.add_spheres(c, r, np.ones((n_spheres, 4)))
    orgs, dirs, target = [], [], []
    for i in range(n_spheres):
        d = _unit(rng.normal(size=(rays_per, 3)))
        e = _perp(d, rng)
        delta = 10.0 ** rng.uniform(-6, -2.5, rays_per)
        sign = np.where(np.arange(rays_per) % 2 == 0, 1.0, -1.0)
        b = r[i] * (1.0 - sign * delta)
        o = c[i].astype(np.float64) - 40.0 * d + b[:, None] * e
        orgs.append(o)
        dirs.append(d)
        target.append(np.full(rays_per, i))
    return sc, rays(np.concatenate(orgs).astype(np.float32), np.concatenate(dirs).astype(np.float32)), \
        np.concatenate(target)


def curve_case(rng, n=30, rays_per=64, open_=False, at_end=False):
    """Cylinders of radius 0.3 .. 1 and length 2 .. 6 at ~1e4 (capped cones,
    or open round curves), rays crossing the axis' middle third
    perpendicularly at b = r (1 -+ delta) from the axis; at_end: rays
    perpendicular to the axis passing b = r (1 -+ delta) beyond the end point
    p1 along the axis (the round curve's end sphere). Returns (scene, rays,
    the index of the cylinder each ray grazes)."""
    p0 = (OFFSET + rng.uniform(-20, 20, (n, 3))).astype(np.float32)
    ax = _unit(rng.normal(size=(n, 3)))
    L = rng.uniform(2.0, 6.0, n)
    p1 = (p0.astype(np.float64) + ax * L[:, None]).astype(np.float32)
    r = rng.uniform(0.3, 1.0, n).astype(np.float32)
    sc = Scene()
    sc.add_cylinders(p0, p1, r, np.ones((n, 4)), open=open_)
    orgs, dirs, target = [], [], []
    for i in range(n):
        a64, b64 = p0[i].astype(np.float64), p1[i].astype(np.float64)
        w = _unit(b64 - a64)
        d = _perp(np.tile(w, (rays_per, 1)), rng)
        e = np.cross(w, d)
        delta = 10.0 ** rng.uniform(-6, -2.5, rays_per)
        sign = np.where(np.arange(rays_per) % 2 == 0, 1.0, -1.0)
        b = r[i] * (1.0 - sign * delta)
        if at_end:
            o = b64 + b[:, None] * w - 40.0 * d
        else:
            h = rng.uniform(0.35, 0.65, rays_per) * np.linalg.norm(b64 - a64)
            o = a64 + h[:, None] * w - 40.0 * d + b[:, None] * e
        orgs.append(o)
        dirs.append(d)
        target.append(np.full(rays_per, i))
    return sc, rays(np.concatenate(orgs).astype(np.float32), np.concatenate(dirs).astype(np.float32)), \
        np.concatenate(target)


def fp32_plain_sphere_hits(sc: Scene, r) -> np.ndarray:
    """Brute-force hit flags of the shifted-frame FP32 sphere test with no
    polish (float32 arithmetic throughout)."""
    f32 = np.float32
    o = r["org"].astype(f32)
    d = r["dir"].astype(f32)
    hit = np.zeros(len(r), bool)
    for c4 in zip(sc.sphere_center.astype(f32), sc.sphere_radius.astype(f32)):
        c, rad = c4
        A = (d * d).sum(-1, dtype=f32)
        ts = ((c - o) * d).sum(-1, dtype=f32) / A
        f = (o + d * ts[:, None]) - c
        B = (f * d).sum(-1, dtype=f32)
        C = (f * f).sum(-1, dtype=f32) - rad * rad
        disc = B * B - A * C
        hit |= disc >= 0
    return hit
