"""Seeded synthetic molecule scenes for the benchmark configurations.

The reference's real CueMol exports are Git-LFS pointers, so every workload
is generated here from a seed:

* :func:`protein_atoms` -- a C-alpha random walk (3.8 A steps, bond angle
  uniform in [90, 120] degrees, uniform dihedral) with 3-5 side atoms per
  residue chained at 1.5 A, element drawn from a 5-entry palette.
* :func:`spheres_scene` -- "2,000 spheres of a synthetic protein", radii from
  {1.2, 1.55, 1.7, 1.8} A, POV Phong (diffuse 0.7, specular 0.3, roughness
  0.01), one distant light plus ambient, hard shadows, 800x600 at 1 spp.
* :func:`ball_and_stick_scene` -- 20,000 atom spheres (r 0.4 A) plus every
  pair closer than 1.9 A drawn as two capped half-bond cylinders (r 0.2 A,
  one per atom colour), AO 16 (max distance 8 A), shadows, linear depth fog
  over 60 A, 1920x1080 with 2x2 supersampling.
"""
from __future__ import annotations

import functools
import math
from typing import Tuple

import numpy as np

from .scene import Camera, DistantLight, Fog, Material, RenderOptions, Scene

# C, N, O, S, H-like colours (linear RGB)
PALETTE = np.array([
    [0.55, 0.55, 0.55, 1.0],
    [0.20, 0.30, 1.00, 1.0],
    [1.00, 0.15, 0.15, 1.0],
    [1.00, 0.85, 0.20, 1.0],
    [0.92, 0.92, 0.92, 1.0],
], np.float32)
RADII = np.array([1.2, 1.55, 1.7, 1.8], np.float32)

LIT = Material(ambient=0.2, diffuse=0.7, specular=0.3, roughness=0.01)


def _cross3(a, b) -> np.ndarray:
    """a x b of two 3-vectors in float64, component for component the
    products and differences np.cross forms (without its per-call overhead)."""
    a0, a1, a2 = float(a[0]), float(a[1]), float(a[2])
    b0, b1, b2 = float(b[0]), float(b[1]), float(b[2])
    return np.array([a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0])


def _perp(v: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    a = (1.0, 0.0, 0.0) if abs(v[0]) < 0.9 else (0.0, 1.0, 0.0)
    p = _cross3(v, a)
    p /= np.linalg.norm(p)
    ang = rng.uniform(0.0, 2.0 * math.pi)
    q = _cross3(v, p)
    return math.cos(ang) * p + math.sin(ang) * q


def protein_atoms(n_atoms: int, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    p, e = _protein_atoms(int(n_atoms), int(seed))
    return p.copy(), e.copy()


@functools.lru_cache(maxsize=8)
def _protein_atoms(n_atoms: int, seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Positions (n, 3) float32 and element indices (n,) of a synthetic chain.

    The C-alpha walk is kept globular like a folded protein: once it leaves a
    sphere holding ~25 A^3 per atom, each step takes the most inward of four
    candidate directions that satisfy the bond-angle constraint. Side atoms
    avoid steric clashes: each takes the first of up to 12 random placements
    that keeps 2.2 A from every atom it is not bonded to (a uniform grid hash
    finds the neighbours), so nearly all sub-1.9 A pairs are real bonds.
    """
    rng = np.random.default_rng(seed)
    pos = []
    elem = []
    ca = np.zeros(3)
    d = np.array([1.0, 0.0, 0.0])
    r_glob = (3.0 * 25.0 * n_atoms / (4.0 * math.pi)) ** (1.0 / 3.0)
    cell = 2.2
    grid: dict = {}

    def key(p):
        return (int(math.floor(p[0] / cell)), int(math.floor(p[1] / cell)), int(math.floor(p[2] / cell)))

    flat = []  # the positions again as Python float triples (fast scalar distance tests)

    def clear(p, skip):
        kx, ky, kz = key(p)
        px, py, pz = float(p[0]), float(p[1]), float(p[2])
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((kx + dx, ky + dy, kz + dz), ()):
                        if j == skip:
                            continue
                        qx, qy, qz = flat[j]
                        ex, ey, ez = qx - px, qy - py, qz - pz
                        # the same float64 sum as np.sum of the squared differences
                        if ex * ex + ey * ey + ez * ez < cell * cell:
                            return False
        return True

    def place(p, e):
        grid.setdefault(key(p), []).append(len(pos))
        pos.append(p)
        flat.append((float(p[0]), float(p[1]), float(p[2])))
        elem.append(e)

    while len(pos) < n_atoms:
        ca_idx = len(pos)
        place(ca.copy(), 0)
        k = int(rng.integers(3, 6))
        prev, prev_idx, prevdir = ca, ca_idx, _perp(d, rng)
        for _ in range(k):
            # chained side atoms, ~tetrahedral turn at each step
            for _try in range(12):
                nd = prevdir * math.cos(math.radians(70.5)) + _perp(prevdir, rng) * math.sin(math.radians(70.5))
                nd /= np.linalg.norm(nd)
                p = prev + 1.5 * nd
                if clear(p, prev_idx):
                    break
            place(p, int(rng.integers(0, 5)))
            prev, prev_idx, prevdir = p, len(pos) - 1, nd
            if len(pos) >= n_atoms:
                break
        # next C-alpha: bond angle theta in [90, 120] deg between successive steps
        theta = math.radians(rng.uniform(90.0, 120.0))
        turn = math.pi - theta
        cands = [d * math.cos(turn) + _perp(d, rng) * math.sin(turn) for _ in range(8)]
        free = [c for c in cands if clear(ca + 3.8 * c, -1)] or cands
        if np.linalg.norm(ca) > r_glob:
            nd = min(free[:4], key=lambda c: float(np.linalg.norm(ca + 3.8 * c)))
        else:
            nd = free[0]
        d = nd / np.linalg.norm(nd)
        ca = ca + 3.8 * d
    p = np.asarray(pos[:n_atoms], np.float32)
    p -= p.mean(axis=0, keepdims=True)
    return p, np.asarray(elem[:n_atoms], np.int32)


def bonds_within(pos: np.ndarray, cutoff: float) -> np.ndarray:
    """All pairs (i < j) with |p_i - p_j| < cutoff, via a uniform grid."""
    cell = np.floor(pos / cutoff).astype(np.int64)
    cell -= cell.min(axis=0)
    dims = cell.max(axis=0) + 1
    key = (cell[:, 0] * dims[1] + cell[:, 1]) * dims[2] + cell[:, 2]
    order = np.argsort(key, kind="stable")
    skey = key[order]
    pairs = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nk = ((cell[:, 0] + dx) * dims[1] + (cell[:, 1] + dy)) * dims[2] + (cell[:, 2] + dz)
                lo = np.searchsorted(skey, nk, "left")
                hi = np.searchsorted(skey, nk, "right")
                cnt = hi - lo
                if cnt.max(initial=0) == 0:
                    continue
                i = np.repeat(np.arange(len(pos)), cnt)
                j = order[np.concatenate([np.arange(a, b) for a, b in zip(lo, hi) if b > a])] if cnt.sum() else np.zeros(0, np.int64)
                m = i < j
                i, j = i[m], j[m]
                dd = np.linalg.norm(pos[i] - pos[j], axis=1)
                sel = dd < cutoff
                pairs.append(np.stack([i[sel], j[sel]], axis=1))
    if not pairs:
        return np.zeros((0, 2), np.int64)
    pr = np.concatenate(pairs)
    pr = np.unique(pr, axis=0)
    return pr


def frame_camera(pos: np.ndarray, pad: np.ndarray, fovy: float = 40.0, aspect: float = 4 / 3) -> Tuple[Camera, float]:
    """Perspective camera on +z looking down -z that frames the bounding sphere."""
    lo = (pos - pad[:, None]).min(axis=0)
    hi = (pos + pad[:, None]).max(axis=0)
    c = 0.5 * (lo + hi)
    r = float(np.max(np.linalg.norm(pos - c, axis=1) + pad))
    half = math.radians(fovy) * 0.5
    half_w = math.atan(math.tan(half) * aspect)
    dist = r / math.sin(min(half, half_w)) * 1.02
    cam = Camera(position=(float(c[0]), float(c[1]), float(c[2] + dist)), direction=(0.0, 0.0, -1.0),
                 up=(0.0, 1.0, 0.0), fovy=fovy)
    return cam, dist


def key_light() -> DistantLight:
    return DistantLight(direction=(-0.35, -0.45, -0.82), color=(1.0, 1.0, 1.0), intensity=0.9)


def spheres_scene(n_atoms: int = 2000, width: int = 800, height: int = 600, seed: int = 0
                  ) -> Tuple[Scene, RenderOptions]:
    """Plain method, small: 2,000 CPK spheres, hard shadows, no AO."""
    rng = np.random.default_rng(seed + 1)
    pos, elem = protein_atoms(n_atoms, seed)
    radii = RADII[rng.integers(0, len(RADII), n_atoms)]
    sc = Scene()
    sc.add_spheres(pos, radii, PALETTE[elem], LIT)
    sc.camera, _ = frame_camera(pos, radii, 40.0, width / height)
    sc.lights = [key_light()]
    sc.ambient_color = (1.0, 1.0, 1.0)
    sc.background = (0.0, 0.0, 0.0)
    opt = RenderOptions(width=width, height=height, supersample=1, shadows=True, ao_samples=0)
    return sc, opt


def ball_and_stick_scene(n_atoms: int = 20000, width: int = 1920, height: int = 1080,
                         supersample: int = 2, ao_samples: int = 16, seed: int = 0
                         ) -> Tuple[Scene, RenderOptions]:
    """Ball-and-stick: atoms r 0.4, half-bond capped cylinders r 0.2, AO, shadows, fog."""
    pos, elem = protein_atoms(n_atoms, seed)
    bonds = bonds_within(pos, 1.9)
    sc = Scene()
    sc.add_spheres(pos, np.full(n_atoms, 0.4, np.float32), PALETTE[elem], LIT)
    if len(bonds):
        i, j = bonds[:, 0], bonds[:, 1]
        mid = 0.5 * (pos[i] + pos[j])
        p0 = np.concatenate([pos[i], mid])
        p1 = np.concatenate([mid, pos[j]])
        col = np.concatenate([PALETTE[elem[i]], PALETTE[elem[j]]])
        # interleave so each bond's two halves are adjacent
        order = np.stack([np.arange(len(bonds)), np.arange(len(bonds)) + len(bonds)], axis=1).reshape(-1)
        sc.add_cylinders(p0[order], p1[order], np.full(2 * len(bonds), 0.2, np.float32), col[order], LIT,
                         open=False)
    sc.camera, dist = frame_camera(pos, np.full(n_atoms, 0.4, np.float32), 40.0, width / height)
    sc.lights = [key_light()]
    sc.ambient_color = (1.0, 1.0, 1.0)
    sc.background = (0.0, 0.0, 0.0)
    sc.fog = Fog(enabled=True, color=(0.0, 0.0, 0.0), start=float(dist), end=float(dist + 60.0))
    opt = RenderOptions(width=width, height=height, supersample=supersample, shadows=True,
                        ao_samples=ao_samples, ao_distance=8.0)
    return sc, opt


def scaled(scene_fn, scale: float, **kw):
    """Same scene at a reduced image size (for oracle-speed parity runs)."""
    sc, opt = scene_fn(**kw)
    opt.width = max(8, int(opt.width * scale))
    opt.height = max(8, int(opt.height * scale))
    return sc, opt


def gaussian_density(pos: np.ndarray, sigma: float = 1.0, spacing: float = 0.5, pad: float = 4.0
                     ) -> Tuple[np.ndarray, np.ndarray]:
    """Sum of unit Gaussians exp(-|x - a|^2 / (2 sigma^2)) on a regular grid.

    Returns (density float32 (nx, ny, nz), origin (3,)). The Gaussian factors
    per axis, so each x slab is one matrix product over the atoms within
    `pad` sigma of it: rho[X] = (gx(X) * Gy)^T Gz, exact up to that cutoff
    along x (no cutoff along y and z).
    """
    p = np.asarray(pos, np.float64)
    lo = p.min(axis=0) - pad * sigma
    hi = p.max(axis=0) + pad * sigma
    dims = np.ceil((hi - lo) / spacing).astype(np.int64) + 1
    axes = [lo[k] + spacing * np.arange(dims[k]) for k in range(3)]
    inv = 1.0 / (2.0 * sigma * sigma)
    gy = np.exp(-(axes[1][None, :] - p[:, 1:2]) ** 2 * inv)  # (atoms, ny)
    gz = np.exp(-(axes[2][None, :] - p[:, 2:3]) ** 2 * inv)  # (atoms, nz)
    order = np.argsort(p[:, 0], kind="stable")
    xs = p[order, 0]
    rho = np.empty((int(dims[0]), int(dims[1]), int(dims[2])), np.float32)
    cut = pad * sigma
    for X in range(int(dims[0])):
        x = axes[0][X]
        a, b = np.searchsorted(xs, x - cut), np.searchsorted(xs, x + cut, "right")
        if a == b:
            rho[X] = 0.0
            continue
        sel = order[a:b]
        gx = np.exp(-(x - p[sel, 0]) ** 2 * inv)
        rho[X] = ((gy[sel] * gx[:, None]).T @ gz[sel]).astype(np.float32)
    return rho, lo


def surface_nets(rho: np.ndarray, iso: float, origin: np.ndarray, spacing: float
                 ) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Isosurface {rho = iso} as an indexed triangle mesh (naive surface nets,
    Gibson 1998): one vertex per grid cell the surface crosses, at the mean of
    the cell's edge crossings; one quad (two triangles) per grid edge with a
    sign change, joining the four cells around it, wound so the triangle
    normal points to decreasing density (outwards). Vertex normals are the
    normalised negative density gradient at the vertex (trilinear blend of
    central differences). Returns (positions (v, 3), normals (v, 3),
    index (t, 3) uint32)."""
    inside = rho > iso
    nx, ny, nz = rho.shape
    cx, cy, cz = nx - 1, ny - 1, nz - 1
    # cell corner signs -> active cells
    corners = [(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    cnt_in = np.zeros((cx, cy, cz), np.int8)
    for (i, j, k) in corners:
        cnt_in += inside[i:i + cx, j:j + cy, k:k + cz]
    active = (cnt_in > 0) & (cnt_in < 8)
    cell_id = np.full((cx, cy, cz), -1, np.int64)
    ai = np.nonzero(active)
    nv = len(ai[0])
    cell_id[ai] = np.arange(nv)
    # vertex = mean of the crossings of the cell's 12 edges
    acc = np.zeros((nv, 3), np.float64)
    num = np.zeros(nv, np.float64)
    base = np.stack(ai, axis=1).astype(np.float64)
    for axis in range(3):
        for a in (0, 1):
            for b in (0, 1):
                off0 = [0, 0, 0]
                o = [k for k in range(3) if k != axis]
                off0[o[0]] = a
                off0[o[1]] = b
                off1 = list(off0)
                off1[axis] = 1
                v0 = rho[ai[0] + off0[0], ai[1] + off0[1], ai[2] + off0[2]].astype(np.float64)
                v1 = rho[ai[0] + off1[0], ai[1] + off1[1], ai[2] + off1[2]].astype(np.float64)
                cross = (v0 > iso) != (v1 > iso)
                t = np.where(cross, (iso - v0) / np.where(cross, v1 - v0, 1.0), 0.0)
                pt = base + np.asarray(off0, np.float64)
                pt[:, axis] += t
                acc += pt * cross[:, None]
                num += cross
    vgrid = acc / num[:, None]
    positions = (origin[None, :] + spacing * vgrid).astype(np.float32)
    # normals: -grad(rho), trilinear from central differences
    g = np.stack(np.gradient(rho.astype(np.float64)), axis=-1)
    i0 = np.clip(np.floor(vgrid).astype(np.int64), 0, np.array([nx - 2, ny - 2, nz - 2]))
    f = vgrid - i0
    grad = np.zeros((nv, 3))
    for (i, j, k) in corners:
        w = (f[:, 0] if i else 1 - f[:, 0]) * (f[:, 1] if j else 1 - f[:, 1]) * (f[:, 2] if k else 1 - f[:, 2])
        grad += w[:, None] * g[i0[:, 0] + i, i0[:, 1] + j, i0[:, 2] + k]
    nrm = -grad
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    # quads around sign-changing grid edges
    tris = []
    for axis in range(3):
        o1, o2 = (axis + 1) % 3, (axis + 2) % 3  # o1 x o2 = +axis
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[axis] = slice(0, rho.shape[axis] - 1)
        sl1[axis] = slice(1, rho.shape[axis])
        a = inside[tuple(sl0)]
        b = inside[tuple(sl1)]
        ch = a != b
        # the four cells around the edge need o1, o2 indices >= 1 and < dim - 1
        for k in (o1, o2):
            s = [slice(None)] * 3
            s[k] = 0
            ch[tuple(s)] = False
            s[k] = ch.shape[k] - 1
            ch[tuple(s)] = False
        e = np.nonzero(ch)
        ei = np.stack(e, axis=1)
        outward = a[e]  # inside at the lower end: density falls along +axis

        def cell(d1, d2):
            c = ei.copy()
            c[:, o1] -= d1
            c[:, o2] -= d2
            return cell_id[c[:, 0], c[:, 1], c[:, 2]]
        q0, q1, q2, q3 = cell(1, 1), cell(0, 1), cell(0, 0), cell(1, 0)
        # q0 -> q1 -> q2 -> q3 turns counter-clockwise about +axis
        t1 = np.stack([q0, q1, q2], axis=1)
        t2 = np.stack([q0, q2, q3], axis=1)
        flip = ~outward
        t1[flip] = t1[flip][:, ::-1]
        t2[flip] = t2[flip][:, ::-1]
        tris.append(np.concatenate([t1, t2]))
    index = np.concatenate(tris).astype(np.uint32)
    return positions, nrm.astype(np.float32), index


@functools.lru_cache(maxsize=2)
def _surface_mesh(n_atoms: int, seed: int, sigma: float, iso: float, spacing: float):
    pos, elem = protein_atoms(n_atoms, seed)
    rho, origin = gaussian_density(pos, sigma, spacing)
    v, nrm, idx = surface_nets(rho, iso, origin, spacing)
    # per-vertex colour of the nearest atom
    from scipy.spatial import cKDTree
    _, near = cKDTree(pos).query(v, k=1)
    return pos, elem, v, nrm, idx, elem[near]


def molecular_surface_scene(n_atoms: int = 40000, width: int = 1920, height: int = 1080,
                            supersample: int = 1, ao_samples: int = 16, alpha: float = 0.6,
                            sigma: float = 1.0, iso: float = 0.5, spacing: float = 0.5, seed: int = 0
                            ) -> Tuple[Scene, RenderOptions]:
    """Molecular surface: the Gaussian-density isosurface (sigma 1 A, iso 0.5,
    0.5 A grid) of a 40,000-atom synthetic protein as an indexed triangle mesh
    with per-vertex normals, coloured by the nearest atom at alpha 0.6, over an
    opaque inner CPK layer of the same 40,000 atoms; shadows, AO 16 (8 A),
    1920x1080 at 1 spp."""
    pos, elem, v, nrm, idx, velem = _surface_mesh(int(n_atoms), int(seed), float(sigma), float(iso),
                                                  float(spacing))
    rng = np.random.default_rng(seed + 1)
    radii = RADII[rng.integers(0, len(RADII), len(pos))]
    col = PALETTE[velem].copy()
    col[:, 3] = alpha
    sc = Scene()
    sc.set_mesh(v, nrm, col, idx, material=LIT)
    sc.add_spheres(pos, radii, PALETTE[elem], LIT)
    sc.camera, _ = frame_camera(v, np.zeros(len(v), np.float32), 40.0, width / height)
    sc.lights = [key_light()]
    sc.ambient_color = (1.0, 1.0, 1.0)
    sc.background = (0.0, 0.0, 0.0)
    opt = RenderOptions(width=width, height=height, supersample=supersample, shadows=True,
                        ao_samples=ao_samples, ao_distance=8.0)
    return sc, opt


def pathtraced_surface_scene(n_atoms: int = 40000, width: int = 1920, height: int = 1080, spp: int = 1,
                             bounces: int = 4, n_lights: int = 64, seed: int = 0, sigma: float = 1.0,
                             iso: float = 0.5, spacing: float = 0.5):
    """Path-traced GI over the molecular surface: the Gaussian isosurface of
    :func:`molecular_surface_scene` (its vertex colours the transmission tint)
    as a GGX dielectric over the CPK spheres of the same atoms, each element
    a GGX conductor (the palette colour its F0) or a GGX dielectric;
    roughness uniform in [0.05, 0.6] per material, index of refraction in
    [1.33, 1.6]; ``n_lights`` spherical area lights (radius 3-8 % of the
    molecule's bounding radius, 1.1-1.6 bounding radii from its centre,
    random warm-to-cool tints), a dim constant environment; 4 bounces.

    Returns (Scene, PtOptions); the scene's camera frames the surface."""
    from .pathtrace import CONDUCTOR, DIELECTRIC, PtMaterial, PtOptions
    pos, elem, v, nrm, idx, velem = _surface_mesh(int(n_atoms), int(seed), float(sigma), float(iso),
                                                  float(spacing))
    rng = np.random.default_rng(seed + 7)
    radii = RADII[np.random.default_rng(seed + 1).integers(0, len(RADII), len(pos))]
    sc = Scene()
    sc.set_mesh(v, nrm, PALETTE[velem], idx, material=LIT)
    sc.add_spheres(pos, radii, PALETTE[elem], LIT)
    # one material index per element (the path tracer's table is indexed by it)
    sc.prim_materials = [Material(ambient=0.2, diffuse=0.7 - 0.01 * k, specular=0.3, roughness=0.01)
                         for k in range(len(PALETTE))]
    sc.sphere_material = np.asarray(elem, np.int32)
    sc.camera, _ = frame_camera(v, np.zeros(len(v), np.float32), 40.0, width / height)
    sc.background = (0.0, 0.0, 0.0)
    rough = lambda: float(rng.uniform(0.05, 0.6))  # noqa: E731
    mesh_mats = [PtMaterial(DIELECTRIC, rough(), float(rng.uniform(1.33, 1.6)))]
    prim_mats = []
    for _ in range(len(PALETTE)):
        if rng.random() < 0.6:
            prim_mats.append(PtMaterial(CONDUCTOR, rough(), 1.5))
        else:
            prim_mats.append(PtMaterial(DIELECTRIC, rough(), float(rng.uniform(1.33, 1.6))))
    lo, hi = v.min(axis=0), v.max(axis=0)
    c = 0.5 * (lo + hi)
    R = float(np.max(np.linalg.norm(v - c, axis=1)))
    d = rng.normal(size=(n_lights, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    centers = c + d * (R * rng.uniform(1.1, 1.6, n_lights))[:, None]
    lr = R * rng.uniform(0.03, 0.08, n_lights)
    tint = np.stack([np.ones(n_lights), rng.uniform(0.7, 1.0, n_lights), rng.uniform(0.4, 1.0, n_lights)], 1)
    tint = np.where(rng.random(n_lights)[:, None] < 0.5, tint, tint[:, ::-1])
    # each light's radiance so that the summed irradiance at the centre is ~1
    E1 = np.pi * (lr / np.linalg.norm(centers - c, axis=1)) ** 2
    Le = tint * (2.0 / (E1.sum() + 1e-12))
    lights = np.concatenate([centers, lr[:, None], Le], 1).astype(np.float32)
    opt = PtOptions(width=width, height=height, spp=spp, bounces=bounces, seed=seed, lights=lights,
                    env=(0.03, 0.03, 0.035), prim_materials=prim_mats, mesh_materials=mesh_mats,
                    restir_candidates=8, restir_spatial=4, restir_radius=16.0)
    return sc, opt
