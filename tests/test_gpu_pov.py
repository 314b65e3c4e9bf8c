"""A scene read from CueMol POV-Ray files (the 1ab0_scene1.pov viewing setup
with a generated geometry include: a mesh2 surface with per-corner texture
ids, ball-and-stick spheres and cylinders, edge_line outlines, two
sections) renders on the gfx950 path as the oracle renders it, under the
per-pixel gates of the benchmark parity tests."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import binding as ob
from test_gpu_parity import _compare, _gpu
from umbreon_amd import capi, pov
from umbreon_amd.scene import RenderOptions

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
CUEMOL = {"_stereo": 0.0, "_iod": 0.03, "_perspective": 0.0, "_shadow": 0.0,
          "_light_inten": 1.3, "_flash_frac": 0.6, "_amb_frac": 0.0}


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


def _include_text(seed: int = 3) -> str:
    rng = np.random.default_rng(seed)
    # a UV sphere of radius 9 as mesh2 with two colours per corner
    nu, nv = 48, 24
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    v = 9.0 * n + np.array([-8.0, 0.0, 0.0])
    faces = []
    for i in range(nv):
        for j in range(nu):
            a, b = i * nu + j, i * nu + (j + 1) % nu
            c, d = a + nu, b + nu
            faces += [(a, c, b), (b, c, d)]
    vs = ", ".join(f"<{x:.6f},{y:.6f},{z:.6f}>" for x, y, z in v)
    ns = ", ".join(f"<{x:.6f},{y:.6f},{z:.6f}>" for x, y, z in n)
    fs = ", ".join(f"<{a},{b},{c}>,{a % 2},{b % 2},{c % 2}" for a, b, c in faces)
    out = ["#declare _s_tex_0 = texture { pigment { color rgb <0.8,0.3,0.2> } finish { ambient 0.2 diffuse 0.7 specular 0.3 roughness 0.02 } }",
           "union {", "#if (_show_34_35)",
           "mesh2 {", f"vertex_vectors {{ {len(v)}, {vs} }}", f"normal_vectors {{ {len(n)}, {ns} }}",
           "texture_list { 2, texture { pigment { color rgb <0.2,0.5,0.9> } }, texture { pigment { color rgb <0.9,0.9,0.3> } } }",
           f"face_indices {{ {len(faces)}, {fs} }}", "}", "#end", "#if (_show_1_2)"]
    atoms = rng.uniform(-6, 6, size=(60, 3)) + np.array([10.0, 0.0, 0.0])
    for p in atoms:
        out.append(f"sphere {{ <{p[0]:.5f},{p[1]:.5f},{p[2]:.5f}>, 0.7 texture {{ _s_tex_0 }} }}")
    for k in range(59):
        a, b = atoms[k], atoms[k + 1]
        out.append(f"cylinder {{ <{a[0]:.5f},{a[1]:.5f},{a[2]:.5f}>, <{b[0]:.5f},{b[1]:.5f},{b[2]:.5f}>, 0.25 "
                   "texture { pigment { color rgb <0.6,0.6,0.6> } finish { diffuse 0.8 } } }")
    for k in range(0, 58, 2):
        a, b = atoms[k], atoms[k + 2]
        out.append(f"edge_line(<{a[0]:.5f},{a[1]:.5f},{a[2]:.5f}>, <0,0,1>, <{b[0]:.5f},{b[1]:.5f},{b[2]:.5f}>, "
                   "<0,0,1>, 0.1, 0.06, _s_tex_0, <0,0,0>)")
    out += ["#end", "}"]
    return "\n".join(out)


@pytest.mark.parametrize("model", ["pov", "principled"])
def test_pov_scene_parity(device, tmp_path, model):
    src = open(os.path.join(HERE, "golden", "pov", "1ab0_scene1.pov")).read()
    (tmp_path / "s.pov").write_text(src.replace("1ab0_scene1.inc", "s.inc"))
    (tmp_path / "s.inc").write_text(_include_text())
    sc, groups, ao_dist = pov.load_pov_scene(str(tmp_path / "s.pov"), 256, 256, CUEMOL, material_model=model)
    if model == "principled":
        assert sc.mesh_material.model == 1 and any(m.model == 1 and m.pbr_specular > 0 for m in sc.prim_materials)
    print(f"pov scene: {sc.n_triangles} triangles, {len(sc.sphere_radius)} spheres, "
          f"{len(sc.cyl_radius)} cylinders, groups {groups}", flush=True)
    assert groups == ["", "_34_35", "_1_2"]
    assert sc.n_triangles == 2 * 48 * 24 and len(sc.sphere_radius) == 60 and len(sc.cyl_radius) == 59 + 29
    opt = RenderOptions(width=256, height=256, supersample=1, shadows=True, ao_samples=8, ao_distance=ao_dist)
    g = _gpu(device, sc, opt, hires_only=True, want_ids=True, want_ao=True)
    c = ob.render(sc, opt, hires_only=True)
    _compare(f"pov scene ({model})", g, c, ao=True)
    hit = g.hit_id != capi.MISS_ID
    assert hit.mean() > 0.1  # the geometry is in view
