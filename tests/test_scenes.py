"""Workload generators: the synthetic protein, bonds and the Gaussian-density
molecular surface (closed, consistently wound, outward normals)."""
from __future__ import annotations

import numpy as np

from umbreon_amd import scenes


def test_gaussian_density_single_atom_closed_form():
    pos = np.array([[0.3, -0.2, 0.1]], np.float32)
    rho, org = scenes.gaussian_density(pos, sigma=1.0, spacing=0.5)
    idx = np.argwhere(rho == rho.max())[0]
    x = org + 0.5 * idx
    # the grid maximum is the grid point nearest the atom, with value exp(-r^2/2)
    r2 = float(np.sum((x - pos[0].astype(np.float64)) ** 2))
    assert abs(float(rho.max()) - np.exp(-r2 / 2.0)) < 1e-6
    # separable sum == direct sum at a few random grid points
    rng = np.random.default_rng(0)
    for _ in range(5):
        i = tuple(int(rng.integers(0, n)) for n in rho.shape)
        p = org + 0.5 * np.array(i)
        assert abs(float(rho[i]) - np.exp(-np.sum((p - pos[0]) ** 2) / 2.0)) < 1e-6


def test_surface_nets_sphere():
    """An isolated atom's iso-0.5 surface is the sphere r = sqrt(2 ln 2)."""
    pos = np.zeros((1, 3), np.float32)
    rho, org = scenes.gaussian_density(pos, sigma=1.0, spacing=0.1)
    v, n, idx = scenes.surface_nets(rho, 0.5, org, 0.1)
    r = np.linalg.norm(v, axis=1)
    r0 = np.sqrt(2.0 * np.log(2.0))
    assert np.abs(r - r0).max() < 0.02
    # normals radial and outward
    assert np.min(np.sum(n * v / r[:, None], axis=1)) > 0.99
    _check_closed_outward(v, n, idx)


def _check_closed_outward(v, n, idx):
    i = idx.astype(np.int64)
    e = np.concatenate([i[:, [0, 1]], i[:, [1, 2]], i[:, [2, 0]]])
    key = e[:, 0] * len(v) + e[:, 1]
    rkey = e[:, 1] * len(v) + e[:, 0]
    assert np.isin(rkey, key).all()  # every edge appears in both directions: closed
    p = v[i]
    tn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    vn = n[i[:, 0]] + n[i[:, 1]] + n[i[:, 2]]
    assert (np.sum(tn * vn, axis=1) > 0).mean() > 0.995  # winding agrees with the outward normals


def test_molecular_surface_scene_small():
    sc, opt = scenes.molecular_surface_scene(n_atoms=600, width=64, height=48)
    assert len(sc.index) > 1000 and len(sc.sphere_center) == 600
    assert np.allclose(sc.colors[:, 3], 0.6)
    _check_closed_outward(sc.positions, sc.normals, sc.index)
    assert sc.index.max() < len(sc.positions)


def test_bonds_within_brute_force():
    pos, _ = scenes.protein_atoms(400, 3)
    b = scenes.bonds_within(pos, 1.9)
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    i, j = np.nonzero(np.triu(d < 1.9, 1))
    assert np.array_equal(b, np.stack([i, j], axis=1))
