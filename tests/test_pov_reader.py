"""The POV scene reader and the mesh2 geometry reader against the reference's
own checks:
  tests/test_pov_scene_reader.cpp:33-130  (synthetic template scene)
  tests/test_pov_scene_reader.cpp:133-197 (a real CueMol .pov with the
      CueMol command-line constants; 1ab0_scene1.pov, its expected zoom 44.533395
      and white background from test_pov_scene_reader.cpp:203-207)
  tests/test_mesh2_reader.cpp:13-120      (synthetic mesh2, colour
      expressions, cylinder open flags and edge tags)
The reference's test1.inc geometry is stored as a git-lfs pointer without
its contents, so the geometry side is pinned on synthetic inputs only.
"""
import math
import os

import numpy as np
import pytest

from umbreon_amd import pov
from umbreon_amd.scene import Material

HERE = os.path.dirname(__file__)

SYNTHETIC = r'''
#version 3.7;
global_settings { assumed_gamma 2.2 }
#declare _bgcolor = <0.1,0.2,0.3>;
background { color rgb _bgcolor }
#ifndef (_perspective)
  #declare _perspective = 1;
#end
#declare _distance = 100.0;
#declare _zoomy = 10.0;
#declare _zoomx = _zoomy * image_width/image_height;
#macro SpecLighting(a,b,c,d)
  light_source { <1,1,1> color rgb c }
#end
#macro FlashLighting(a)
  light_source { <0,0,1> color rgb a }
#end
camera {
 #if (_perspective)
 perspective
 location <0,0,_distance>
 angle 30
 #else
 orthographic
 direction <0,0,-1>
 up <0, _zoomy, 0>
 right <_zoomx, 0, 0>
 location <0,0,_distance>
 look_at <0,0,0>
 #end
}
SpecLighting(1, _distance, 0.5, 1)
FlashLighting(0.75)
fog {
  distance 30.0/2
  color rgbf <0.1,0.2,0.3,0>
  fog_type 2
  fog_offset 0
  fog_alt 1.0e-10
  up <0,0,1>
}
#declare _scene = #include "geom.inc"
object { _scene }
'''

CUEMOL = {"_stereo": 0.0, "_iod": 0.03, "_perspective": 0.0, "_shadow": 0.0,
          "_light_inten": 1.3, "_flash_frac": 0.6, "_amb_frac": 0.0}
INV3 = -1.0 / math.sqrt(3.0)


def test_tokenizer():
    t = pov.tokenize('a_1 = .5e+2 /* x\n y */ "s t" // c\n <1,2>')
    kinds = [(x.kind, x.s, x.num) for x in t]
    assert kinds[0] == (pov.IDENT, "a_1", 0.0)
    assert kinds[1] == (pov.PUNCT, "=", 0.0)
    assert kinds[2] == (pov.NUM, "", 50.0)
    assert kinds[3] == (pov.STR, "s t", 0.0)
    assert [k[1] for k in kinds[4:9]] == ["<", "", ",", "", ">"]
    assert t[3].line == 2 and t[-1].kind == pov.END and t[-1].line == 3
    # '1e' is a number followed by an identifier; '.x' is punctuation
    t = pov.tokenize("1e x .y")
    assert [(x.kind, x.s) for x in t[:4]] == [(pov.NUM, ""), (pov.IDENT, "e"), (pov.IDENT, "x"), (pov.PUNCT, ".")]


def test_values():
    assert pov.bin_op((2.0,), (3.0,), "*") == (6.0,)
    assert pov.bin_op((1.0, 2.0, 3.0), (2.0,), "*") == (2.0, 4.0, 6.0)
    assert pov.bin_op((1.0, 2.0), (1.0, 1.0, 1.0), "+") == (2.0, 3.0, 1.0)
    assert pov.bin_op((1.0,), (0.0,), "/") == (0.0,)
    assert pov.comp((), 2) == 0.0 and pov.comp((5.0,), 2) == 5.0 and pov.comp((1.0, 2.0), 2) == 0.0


def test_synthetic_scene():
    r = pov.read_pov_scene_string(SYNTHETIC, 200, 100, {"_perspective": 0.0})
    assert r.camera.orthographic  # the predefined constant survives the #ifndef
    assert r.camera.height == pytest.approx(10.0, abs=1e-4)
    assert r.camera.position == pytest.approx((0, 0, 100), abs=1e-4)
    assert r.camera.direction == pytest.approx((0, 0, -1), abs=1e-4)
    assert r.background == pytest.approx((0.1, 0.2, 0.3), abs=1e-4)
    assert len(r.lights) == 2
    assert r.lights[0].intensity == pytest.approx(0.5, abs=1e-4)
    assert r.lights[0].direction == pytest.approx((INV3, INV3, INV3), abs=1e-3)
    assert r.lights[0].casts_highlight  # shadow argument 1
    assert r.lights[1].intensity == pytest.approx(0.75, abs=1e-4)
    assert not r.lights[1].casts_highlight
    assert r.include_path == "geom.inc"
    f = r.fog
    assert f.enabled and f.type == 2
    assert f.distance == pytest.approx(15.0, abs=1e-4)
    assert f.color == pytest.approx((0.1, 0.2, 0.3), abs=1e-4)
    assert f.up == pytest.approx((0, 0, 1), abs=1e-4)
    assert f.start == pytest.approx(100.0, abs=1e-4)
    assert f.end == pytest.approx(122.5, abs=1e-3)
    assert r.assumed_gamma == pytest.approx(2.2, abs=1e-4)
    # _zoomx = _zoomy * image_width / image_height
    assert r.declares["_zoomx"] == [20.0]
    assert r.declares["image_width"] == [200.0]


def test_synthetic_scene_perspective_branch():
    """Without the predefined constant the #ifndef default selects the
    perspective branch: angle 30 with |up| = |right| = 1 gives fovy 30."""
    r = pov.read_pov_scene_string(SYNTHETIC, 200, 100)
    assert not r.camera.orthographic
    assert r.camera.fovy == pytest.approx(30.0, abs=1e-4)
    # no look_at in that branch: the direction keyword's default -z
    assert r.camera.direction == pytest.approx((0, 0, -1), abs=1e-6)


def test_perspective_fovy_closed_form():
    """fovy = degrees(2 atan(tan(radians(angle) / 2) |up| / |right|))."""
    src = "camera { perspective up <0,1,0>*3/4 right <1,0,0> angle 50 location <0,0,9> look_at <0,0,0> }"
    r = pov.read_pov_scene_string(src)
    want = math.degrees(2.0 * math.atan(math.tan(math.radians(50.0) / 2.0) * 0.75))
    assert r.camera.fovy == pytest.approx(want, abs=1e-4)
    assert r.camera.up == pytest.approx((0, 1, 0), abs=1e-6)


def test_functions_and_conditionals():
    src = """
    #declare A = vnormalize(<3,0,4>);
    #declare B = vdot(<1,2,3>, <4,5,6>);
    #declare C = vcross(<1,0,0>, <0,1,0>);
    #declare D = 2*degrees(atan2(1, 1));
    #declare E = sqrt(-4) + abs(-2) + floor(2.7) + int(-2.7) + min(3,4) + max(3,4) + pow(2,3);
    #declare F = vlength(<3,4,0>);
    #if (0)
      #declare G = 1;
    #else
      #declare G = 2;
    #end
    #ifdef (G)
      #if (1)
        #declare H = 5;
      #end
    #end
    #ifdef (NOPE)
      #declare I = 1;
    #end
    #declare J = unknown_fn(1, 2) + unknown_var + 7;
    """
    d = pov.read_pov_scene_string(src).declares
    assert d["A"] == pytest.approx([0.6, 0.0, 0.8])
    assert d["B"] == [32.0]
    assert d["C"] == [0.0, 0.0, 1.0]
    assert d["D"] == pytest.approx([90.0])
    assert d["E"] == pytest.approx([0 + 2 + 2 - 2 + 3 + 4 + 8])
    assert d["F"] == pytest.approx([5.0])
    assert d["G"] == [2.0] and d["H"] == [5.0] and "I" not in d
    assert d["J"] == [7.0]


def test_spec_lighting_spread():
    """spread > 1: highlights and angular radius atan(spread / 40); spread 1
    with shadow 0: no highlight."""
    r = pov.read_pov_scene_string("SpecLighting(4, 200, 0.3, 0) SpecLighting(1, 200, 0.3, 0)")
    assert r.lights[0].casts_highlight
    assert r.lights[0].angular_radius == pytest.approx(math.atan(0.1), abs=1e-7)
    assert not r.lights[1].casts_highlight and r.lights[1].angular_radius == 0.0


def test_fog_distance_without_declare():
    """Fog start falls back to |look_at - location|, clamped to >= 1."""
    src = "camera { location <0,0,50> look_at <0,0,10> } fog { distance 4 color rgb <1,1,1> }"
    f = pov.read_pov_scene_string(src).fog
    assert f.start == pytest.approx(40.0) and f.end == pytest.approx(46.0)
    f = pov.read_pov_scene_string("fog { distance 2 }").fog
    assert f.start == 200.0 and f.end == 203.0


def test_real_cuemol_scene():
    """1ab0_scene1.pov with the CueMol command-line constants
    (test_pov_scene_reader.cpp:133-197)."""
    r = pov.read_pov_scene(os.path.join(HERE, "golden", "pov", "1ab0_scene1.pov"), 300, 300, CUEMOL)
    assert r.camera.orthographic
    assert r.camera.height == pytest.approx(44.533395, abs=1e-2)
    assert r.camera.position == pytest.approx((0, 0, 200), abs=1e-3)
    assert r.camera.direction == pytest.approx((0, 0, -1), abs=1e-3)
    assert r.background == pytest.approx((1, 1, 1), abs=1e-3)
    assert len(r.lights) == 2
    # key: 1.3 (1 - 0) (1 - 0.6) = 0.52 along -<1,1,1>; flash: 1.3 * 0.6 = 0.78 along -z
    assert r.lights[0].intensity == pytest.approx(0.52, abs=1e-3)
    assert r.lights[0].direction == pytest.approx((INV3, INV3, INV3), abs=1e-3)
    assert r.lights[1].intensity == pytest.approx(0.78, abs=1e-3)
    assert r.lights[1].direction == pytest.approx((0, 0, -1), abs=1e-3)
    f = r.fog
    assert f.enabled and f.type == 2
    assert f.distance == pytest.approx(50.0 / 3.0, abs=1e-2)
    assert f.up == pytest.approx((0, 0, 1), abs=1e-4)
    assert f.color == pytest.approx((1, 1, 1), abs=1e-3)
    assert f.start == pytest.approx(200.0, abs=1e-2)
    assert f.end == pytest.approx(225.0, abs=1e-1)
    assert r.assumed_gamma == pytest.approx(2.2, abs=1e-3)
    assert r.include_path == "1ab0_scene1.inc"
    assert not r.radiosity
    # the file's own defaults apply where nothing is predefined
    assert r.declares["_light_spread"] == [1.0] and r.declares["_show_34_35"] == [1.0]
    # _radiosity defined: the radiosity branch's sphere and no fog change
    rr = pov.read_pov_scene(os.path.join(HERE, "golden", "pov", "1ab0_scene1.pov"), 300, 300,
                            {**CUEMOL, "_radiosity": 1.0, "_transpbg": 1.0})
    assert rr.radiosity and not rr.fog.enabled


MESH2_SYNTHETIC = r'''
  #declare C0 = <1, 0, 0>;
  #declare C1 = <0, 0, 1>;
  // skipped entirely: its #local texture finish is not the mesh material
  #macro junk(a)
    #local tex0 = texture { finish { ambient 1.0 diffuse 0.111 }
                            pigment { color rgb a } }
  #end
  union {
    mesh2 {
      vertex_vectors { 4, <0,0,0>, <1,0,0>, <0,1,0>, <1,1,0> }
      normal_vectors { 4, <0,0,1>, <0,0,1>, <0,0,1>, <0,0,1> }
      texture_list {
        2,
        texture { pigment { color rgb C0 } },
        texture { pigment { color rgb C1 } }
      }
      face_indices { 2, <0,1,2>,0,0,0, <1,3,2>,1,1,1 }
    }
  }
'''


def test_mesh2_synthetic():
    g = pov.read_geometry_string(MESH2_SYNTHETIC)
    assert g.n_triangles == 2
    # v1 and v2 are shared but carry different colours: 6 vertices
    assert len(g.positions) == 6 and g.index.size == 6
    assert tuple(g.colors[g.index[0, 0]][:3]) == (1.0, 0.0, 0.0)
    assert tuple(g.colors[g.index[1, 0]][:3]) == (0.0, 0.0, 1.0)
    assert abs(g.material.diffuse - 0.8) < 1e-4
    # position classes: 4 distinct positions over the 6 vertices
    assert g.pos_class.tolist() == [0, 1, 2, 1, 3, 2]


def test_mesh2_colour_expression():
    src = r'''
      #declare A = <1, 1, 1>;
      #declare B = <0, 0, 0>;
      mesh2 {
        vertex_vectors { 3, <0,0,0>, <1,0,0>, <0,1,0> }
        normal_vectors { 3, <0,0,1>, <0,0,1>, <0,0,1> }
        texture_list { 1, texture { pigment { color rgbt A*0.25+B*0.75 } } }
        face_indices { 1, <0,1,2>, 0, 0, 0 }
      }
    '''
    g = pov.read_geometry_string(src)
    c = g.colors[g.index[0, 0]]
    assert abs(c[0] - 0.25) < 1e-5 and abs(c[2] - 0.25) < 1e-5
    assert c[3] == 1.0  # three components: no transparency term


def test_cylinder_open_flags_and_edge_tags():
    src = r'''
      #declare _tex0 = texture { pigment { color rgb <0,0,0> } }
      edge_line(<0,0,0>, <0,0,1>, <1,0,0>, <0,0,1>, 0.0, 0.05, _tex0, <0,0,0>)
      edge_line2(<0,0,0>, <0,0,1>, 0.0, <1,0,0>, <0,0,1>, 0.0, 0.0, 0.05,
                 _tex0, <0,0,0>)
      cylinder { <0,0,0>, <1,0,0>, 0.05
                 texture { pigment { color rgb <1,0,0> } } }
      cylinder { <0,0,0>, <1,0,0>, 0.05 open
                 texture { pigment { color rgb <0,1,0> } } }
    '''
    g = pov.read_geometry_string(src)
    assert len(g.cylinders) == 4
    assert [c.open for c in g.cylinders] == [True, True, False, True]
    assert [c.from_edge for c in g.cylinders] == [True, True, False, False]
    assert g.cylinders[2].color == (1.0, 0.0, 0.0, 1.0)
    assert g.cylinders[2].material == Material.flat_outline()


def test_edge_line_geometry():
    """edge_line raises both ends along their normals; edge_line2 turns the
    end transparencies into opacities 1 - t clamped to [0, 1]."""
    src = r'''
      edge_line(<0,0,0>, <0,0,1>, <1,0,0>, <0,1,0>, 0.5, 0.1, T, <0.2,0.3,0.4>)
      edge_line2(<0,0,0>, <0,0,1>, 0.25, <1,0,0>, <0,0,1>, 1.5, 0.5, 0.2, T, <1,1,1>)
      edge_line(<0,0,0>, <0,0,1>, <1,0,0>, <0,1,0>, 0.5, 0.0, T, <1,1,1>)
      edge_line(<0,0,0>, <0,0,1>)
    '''
    g = pov.read_geometry_string(src)
    assert len(g.cylinders) == 2  # zero width and short argument lists are dropped
    a, b = g.cylinders
    assert a.p0 == (0.0, 0.0, 0.5) and a.p1 == (1.0, 0.5, 0.0)
    assert a.radius == pytest.approx(0.1) and a.color == pytest.approx((0.2, 0.3, 0.4, 1.0))
    assert a.opacity1 == -1.0
    assert b.p0 == (0.0, 0.0, 0.5) and b.p1 == (1.0, 0.0, 0.5)
    assert b.color[3] == 0.75 and b.opacity1 == 0.0


def test_spheres_textures_and_finishes():
    src = r'''
      #declare _t = texture { pigment { color rgb <0.1,0.2,0.3> } finish { ambient 0.3 diffuse 0.6 specular 0.4 roughness 0.01 metallic } }
      #declare _u = texture { _t pigment { color srgbt <0.5,0.5,0.5,0.25> } }
      sphere { <1,2,3>, 0.5 texture { _t } }
      sphere { <0,0,0>, 2 texture { _u } }
      sphere { <0,0,0>, 1 texture { _x_sl_tex } }
      sphere { <0,0,0>, 0 texture { _t } }
      sphere { <0,0,0>, 1 texture { pigment { color rgb 0.5 } finish { F_MetalC } } }
      sphere { <0,0,0>, 1 }
    '''
    g = pov.read_geometry_string(src)
    s = g.spheres
    assert len(s) == 5  # the zero-radius sphere is dropped
    assert s[0].center == (1.0, 2.0, 3.0) and s[0].radius == 0.5
    assert s[0].color == pytest.approx((0.1, 0.2, 0.3, 1.0))
    m = s[0].material
    assert (m.ambient, m.diffuse, m.specular, m.metallic) == pytest.approx((0.3, 0.6, 0.4, True))
    # _u: its own pigment (sRGB decoded, opacity 1 - t), the base texture's finish
    lin = ((0.5 + 0.055) / 1.055) ** 2.4
    assert s[1].color == pytest.approx((lin, lin, lin, 0.75), abs=1e-6)
    assert s[1].material == m
    # an undeclared texture name: white, flat outline; tagged as an edge joint
    assert s[2].from_edge and s[2].color == (1.0, 1.0, 1.0, 1.0)
    assert s[2].material == Material.flat_outline()
    assert s[3].color == (0.5, 0.5, 0.5, 1.0)
    assert s[3].material.metallic and s[3].material.brilliance == 4.0
    assert s[3].material.roughness == pytest.approx(1.0 / 80.0)
    # no texture at all: black
    assert s[4].color == (0.0, 0.0, 0.0, 1.0)


def test_sections_blocks_and_flat_normals():
    src = r'''
      #declare _f = texture { finish { diffuse 0.5 } }
      mesh2 { vertex_vectors { 3, <0,0,0>, <1,0,0>, <0,1,0> } face_indices { 1, <0,1,2> } }
      #if (_show_12_13)
        mesh2 {
          vertex_vectors { 3, <0,0,0>, <0,0,1>, <0,1,0> }
          texture_list { 1, texture { pigment { rgb <0,1,0> } finish { diffuse 0.25 } } }
          face_indices { 1, <0,1,2>, 0 }
        }
        #if (1) sphere { <0,0,0>, 1 } #end
        cylinder { <0,0,0>, <1,0,0>, 0.1 }
      #end
      #if (_show_7)
        sphere { <0,0,0>, 1 }
      #end
      #ifdef (_show_12_13)
        sphere { <0,0,0>, 1 }
      #end
      sphere { <0,0,0>, 1 }
    '''
    g = pov.read_geometry_string(src)
    assert g.group_names == ["", "_12_13", "_7"]
    assert g.tri_group.tolist() == [0, 1]
    assert [s.group for s in g.spheres] == [1, 2, 1, 0]
    assert g.cylinders[0].group == 1
    # block materials: the first finish seen before each block
    assert g.tri_material.tolist() == [0, 1]
    assert [m.diffuse for m in g.materials] == [0.5, 0.25]
    # no normals: flat face normals; no texture list: white
    assert g.normals[g.index[0, 0]].tolist() == [0.0, 0.0, 1.0]
    assert g.normals[g.index[1, 0]].tolist() == [-1.0, 0.0, 0.0]
    assert g.colors[g.index[0, 0]].tolist() == [1.0, 1.0, 1.0, 1.0]
    assert g.colors[g.index[1, 0]].tolist() == [0.0, 1.0, 0.0, 1.0]
    # position (0,0,0) is shared by both blocks: one class
    assert g.pos_class[g.index[0, 0]] == g.pos_class[g.index[1, 0]]


def test_weld_across_colours_and_per_corner_textures():
    src = r'''
      mesh2 {
        vertex_vectors { 4, <0,0,0>, <1,0,0>, <0,1,0>, <1,1,0> }
        normal_vectors { 4, <0,0,1>, <0,0,1>, <0,0,1>, <0,0,1> }
        texture_list { 2, texture { pigment { rgb 1 } }, texture { pigment { rgb 0 } } }
        face_indices { 2, <0,1,2>, 0, 1, 0, <1,3,2>, 1, 0, 0 }
      }
    '''
    g = pov.read_geometry_string(src)
    # corners (pos, colour): (0,w)(1,k)(2,w) | (1,k)(3,w)(2,w): 4 vertices
    assert len(g.positions) == 4
    assert g.index.tolist() == [[0, 1, 2], [1, 3, 2]]
    assert g.pos_class.tolist() == [0, 1, 2, 3]


def test_geometry_errors():
    with pytest.raises(pov.PovError, match="no geometry"):
        pov.read_geometry_string("#declare a = 1;")
    with pytest.raises(pov.PovError, match="out of range"):
        pov.read_geometry_string("mesh2 { vertex_vectors { 1, <0,0,0> } face_indices { 1, <0,1,2> } }")
    with pytest.raises(pov.PovError, match="unbalanced"):
        pov.read_geometry_string("mesh2 { vertex_vectors { 1, <0,0,0> }")
    with pytest.raises(pov.PovError, match="unbalanced"):
        pov.read_pov_scene_string("camera { orthographic")


def test_load_pov_scene(tmp_path):
    """A .pov plus its include as a Scene: the reader's camera, lights,
    background and fog, the geometry, gamma 2.2, the GI energy balance."""
    src = open(os.path.join(HERE, "golden", "pov", "1ab0_scene1.pov")).read()
    (tmp_path / "s.pov").write_text(src.replace("1ab0_scene1.inc", "s.inc"))
    (tmp_path / "s.inc").write_text(r'''
      union {
      #if (_show_34_35)
      mesh2 { vertex_vectors { 3, <-1,-1,0>, <1,-1,0>, <0,1,0> } face_indices { 1, <0,1,2> } }
      sphere { <0,0,0>, 1 texture { pigment { color rgb <1,0,0> } } }
      cylinder { <0,0,0>, <3,4,0>, 0.5 }
      #end
      }
    ''')
    sc, groups, ao = pov.load_pov_scene(str(tmp_path / "s.pov"), 64, 48, CUEMOL)
    assert groups == ["", "_34_35"]
    assert sc.n_triangles == 1 and len(sc.sphere_radius) == 1 and len(sc.cyl_radius) == 1
    assert sc.camera.orthographic and sc.camera.height == pytest.approx(44.533395, abs=1e-3)
    assert [L.intensity for L in sc.lights] == pytest.approx([0.52, 0.78], abs=1e-3)
    assert sc.fog.enabled and sc.fog.start == pytest.approx(200.0) and sc.fog.end == pytest.approx(225.0, abs=1e-3)
    assert sc.assumed_gamma == 2.2 and sc.ambient_color == (1.0, 1.0, 1.0)
    # bounds (-1,-1,-1)..(3,4,1): diagonal sqrt(16 + 25 + 4)
    assert ao == pytest.approx(0.7 * math.sqrt(45.0), rel=1e-6)
    assert int(sc.sphere_group[0]) == 1 and int(sc.cyl_group[0]) == 1
    # GI: radiosity balance replaces the non-radiosity defaults
    sg, _, _ = pov.load_pov_scene(str(tmp_path / "s.pov"), 64, 48, CUEMOL, gi=True)
    assert sg.ambient_color == pytest.approx((0.8, 0.8, 0.8))
    assert [L.intensity for L in sg.lights] == pytest.approx([1.6 * 0.5 * 0.5, 1.6 * 0.5 * 0.5], abs=1e-6)
    with pytest.raises(pov.PovError, match="_scene"):
        (tmp_path / "n.pov").write_text("camera { }")
        pov.load_pov_scene(str(tmp_path / "n.pov"))


def _exp_to_rough(e: float) -> float:
    return math.sqrt(min(math.sqrt(2.0 / (e + 2.0)), 1.0))


def test_material_convert():
    """The rows of test_material_convert.cpp:27-161."""
    conv = pov.to_principled_material
    out = conv(Material(ambient=0.15, emission=2.0))
    assert out.model == 1 and (out.ambient, out.diffuse, out.emission) == pytest.approx((0.15, 0.8, 2.0))
    assert out.pbr_specular == 0.0 and out.pbr_metallic == 0.0
    out = conv(Material(specular=0.4, roughness=0.02))  # Blinn exponent 50
    assert out.pbr_specular == pytest.approx(0.4) and out.pbr_metallic == 0.0
    assert out.pbr_roughness == pytest.approx(_exp_to_rough(50.0), abs=1e-4)
    out = conv(Material(phong=0.8, phong_size=50.0))
    assert out.pbr_specular == pytest.approx(0.8)
    assert out.pbr_roughness == pytest.approx(_exp_to_rough(50.0), abs=1e-4)
    toon = Material(ambient=0.3, diffuse=0.5, brilliance=0.0, phong=10000.0, phong_size=50.0)
    assert pov.is_non_physical_finish(toon) and conv(toon) == toon
    assert conv(Material(brilliance=0.0)).model == 0
    assert conv(Material(specular=4.0)).model == 0
    unlit = Material(ambient=1.0, diffuse=0.0, specular=0.0)
    assert pov.is_non_physical_finish(unlit) and conv(unlit).model == 0
    assert conv(Material(metallic=True, brilliance=2.0, specular=0.8, roughness=0.05)).model == 1
    assert conv(Material(metallic=True, specular=0.8, roughness=0.05)).pbr_metallic == 1.0
    out = conv(Material(reflection=0.8))
    assert (out.pbr_metallic, out.pbr_roughness, out.reflection) == pytest.approx((1.0, 0.0, 0.8))
    out = conv(Material(reflection=0.3, specular=0.6, roughness=0.02))
    assert out.pbr_metallic == pytest.approx(0.3) and out.pbr_specular == pytest.approx(0.6)
    assert out.pbr_roughness == pytest.approx(_exp_to_rough(50.0), abs=1e-4)
    p = Material(model=1, pbr_metallic=0.7, pbr_roughness=0.25)
    assert conv(p) == p
    # the Blinn exponent is clamped to [1, 1e6]: roughness 0 -> e = 1e6, 2 -> e = 1
    assert conv(Material(specular=0.5, roughness=0.0)).pbr_roughness == pytest.approx(_exp_to_rough(1.0e6), rel=1e-5)
    assert conv(Material(specular=0.5, roughness=2.0)).pbr_roughness == pytest.approx(_exp_to_rough(1.0), rel=1e-6)


def test_load_pov_scene_principled(tmp_path):
    """--material principled: every finish but the edge decoration's."""
    src = open(os.path.join(HERE, "golden", "pov", "1ab0_scene1.pov")).read()
    (tmp_path / "s.pov").write_text(src.replace("1ab0_scene1.inc", "s.inc"))
    (tmp_path / "s.inc").write_text(r'''
      #declare _t = texture { pigment { rgb 1 } finish { specular 0.4 roughness 0.02 } }
      mesh2 { vertex_vectors { 3, <-1,-1,0>, <1,-1,0>, <0,1,0> } texture_list { 1, texture { _t } } face_indices { 1, <0,1,2>, 0 } }
      sphere { <0,0,0>, 1 texture { _t } }
      sphere { <0,0,0>, 1 texture { _x_sl_tex } }
      edge_line(<0,0,0>, <0,0,1>, <1,0,0>, <0,0,1>, 0.0, 0.05, _t, <0,0,0>)
    ''')
    sc, _, _ = pov.load_pov_scene(str(tmp_path / "s.pov"), 64, 48, CUEMOL, material_model="principled")
    assert sc.mesh_material.model == 1
    mats = [sc.prim_materials[i] for i in sc.sphere_material] + [sc.prim_materials[i] for i in sc.cyl_material]
    assert [m.model for m in mats] == [1, 0, 0]  # the real sphere converts; edge joint and edge line stay
    with pytest.raises(ValueError):
        pov.load_pov_scene(str(tmp_path / "s.pov"), material_model="toon")
