// Closed-form checks of the public C++ API on the GPU backend (run by
// tests/test_gpu_cpp_api.py). Restates reference assertions:
//   test_render_shading.cpp:19-36  (quad centre 0.30/0.36/0.42, depth 10)
//   test_render_shading.cpp:39-53  (background on miss)
//   test_render_transparency.cpp:73-229 (group-alpha blend in the sRGB domain)
//   test_render_async.cpp           (async == blocking, progress, pre-cancel)
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>

#include "postprocess/image_ops.hpp"
#include "umbreon.hpp"

namespace {
int failures = 0;
void check(const char* what, bool ok) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}
bool near(float a, float b, float e) { return std::fabs(a - b) <= e; }

umbreon::Mesh quad(umbreon::Vec4 c, float z) {
  umbreon::Mesh m;
  const umbreon::Vec3 p[6] = {{-2, -2, z}, {2, -2, z}, {2, 2, z}, {-2, -2, z}, {2, 2, z}, {-2, 2, z}};
  for (auto& v : p) {
    m.positions.push_back(v);
    m.normals.push_back({0, 0, 1});
    m.colors.push_back(c);
  }
  m.material.ambient = 0.2f;
  m.material.diffuse = 0.8f;
  return m;
}
umbreon::Camera orthoCam() {
  umbreon::Camera c;
  c.position = {0, 0, 10};
  c.direction = {0, 0, -1};
  c.up = {0, 1, 0};
  c.orthographic = true;
  c.height = 4.0f;
  return c;
}
umbreon::DistantLight key() {
  umbreon::DistantLight l;
  l.direction = {0, 0, -1};
  l.intensity = 0.5f;
  return l;
}
constexpr std::size_t kC = (2 * 5 + 2) * 4;
}  // namespace

int main() {
  {
    // render_progress.cpp:109-121: the phase names a UI shows
    using P = umbreon::RenderPhase;
    const std::pair<P, const char*> names[] = {{P::Idle, "Idle"},       {P::Setup, "Setup"},
                                               {P::CoarseAo, "CoarseAO"}, {P::Primary, "Primary"},
                                               {P::GlobalIllum, "GI"},    {P::Edges, "Edges"},
                                               {P::Postprocess, "Postprocess"}, {P::Done, "Done"}};
    bool all = true;
    for (const auto& n : names) all = all && std::string(umbreon::toString(n.first)) == n.second;
    check("toString(RenderPhase) names", all);
  }
  umbreon::RenderOptions o;
  o.width = 5;
  o.height = 5;
  {
    umbreon::Scene s;
    s.mesh = quad({0.5f, 0.6f, 0.7f, 1.0f}, 0.0f);
    s.camera = orthoCam();
    s.lights.push_back(key());
    s.background = {0.4f, 0.1f, 0.2f};
    umbreon::FrameResult f = umbreon::render(s, o);
    check("quad centre R 0.30", near(f.color[kC], 0.30f, 1e-4f));
    check("quad centre G 0.36", near(f.color[kC + 1], 0.36f, 1e-4f));
    check("quad centre B 0.42", near(f.color[kC + 2], 0.42f, 1e-4f));
    check("quad depth 10", near(f.depth[12], 10.0f, 1e-3f));
    umbreon::Scene e;
    e.camera = orthoCam();
    e.background = {0.4f, 0.1f, 0.2f};
    umbreon::FrameResult g = umbreon::render(e, o);
    check("miss -> background", near(g.color[kC], 0.4f, 1e-6f) && near(g.color[kC + 2], 0.2f, 1e-6f));
  }
  {
    // group-alpha: opaque red (group 0, z=0) behind opaque blue (group 1, z=1),
    // blend {group 1, 0.5}: each channel = dec(0.5 enc(1) + 0.5 enc(0))
    umbreon::Scene s;
    umbreon::Mesh m = quad({1, 0, 0, 1}, 0.0f);
    umbreon::Mesh b = quad({0, 0, 1, 1}, 1.0f);
    for (std::size_t i = 0; i < b.positions.size(); ++i) {
      m.positions.push_back(b.positions[i]);
      m.normals.push_back(b.normals[i]);
      m.colors.push_back(b.colors[i]);
    }
    m.triGroupId = {0, 0, 1, 1};
    m.material = umbreon::Material::flatOutline();
    s.mesh = m;
    s.camera = orthoCam();
    s.groupBlend.push_back({1, 0.5f});
    umbreon::FrameResult f = umbreon::render(s, o);
    const float want = umbreon::srgbDecodeF(0.5f);
    check("group blend R", near(f.color[kC], want, 1e-4f));
    check("group blend B", near(f.color[kC + 2], want, 1e-4f));
    check("group blend alpha 1", near(f.color[kC + 3], 1.0f, 1e-6f));
  }
  {
    umbreon::Scene s;
    s.mesh = quad({0.5f, 0.6f, 0.7f, 1.0f}, 0.0f);
    s.camera = orthoCam();
    s.lights.push_back(key());
    umbreon::RenderOptions a = o;
    a.aoSamples = 16;
    a.aoDistance = 100.0f;
    umbreon::FrameResult ref = umbreon::render(s, a);
    umbreon::RenderTask t = umbreon::renderAsync(s, a);
    umbreon::FrameResult got = t.get();
    check("async == blocking", got.color == ref.color);
    check("async task done", t.done());
    umbreon::RenderProgress p;
    umbreon::FrameResult f = umbreon::render(s, a, p);
    check("progress reaches 1", near(p.fraction(), 1.0f, 1e-4f) && p.phase() == umbreon::RenderPhase::Done);
    umbreon::RenderProgress pc;
    pc.requestCancel();
    umbreon::FrameResult fc = umbreon::render(s, a, pc);
    check("pre-cancelled render is cancelled", fc.cancelled && pc.phase() != umbreon::RenderPhase::Done);
  }
  {
    umbreon::Scene s;
    s.camera = orthoCam();
    umbreon::RenderOptions g = o;
    g.gi = true;
    g.giIntegrator = 0;  // the irradiance cache: not implemented
    bool threw = false;
    try {
      umbreon::render(s, g);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    check("unsupported option throws", threw);
    // the default GI (pt2, denoised) renders
    umbreon::RenderOptions d = o;
    d.gi = true;
    d.pt1Spp = 2;
    bool ok = true;
    try {
      const umbreon::FrameResult f = umbreon::render(s, d);
      ok = f.pt1DenoiserUsed == 1 && f.indirect.size() == f.depth.size() * 3;
    } catch (const std::exception&) {
      ok = false;
    }
    check("default GI renders", ok);
  }
  {
    // a principled material with metallic 0 and specular 0 is the POV
    // diffuse-only finish bit for bit (test_principled.cpp:165-199); a metal
    // is not
    umbreon::Scene s;
    s.mesh = quad({0.5f, 0.6f, 0.7f, 1.0f}, 0.0f);
    s.camera = orthoCam();
    s.lights.push_back(key());
    s.background = {0.4f, 0.1f, 0.2f};
    const umbreon::FrameResult pov = umbreon::render(s, o);
    s.mesh.material.model = umbreon::ShadingModel::Principled;
    s.mesh.material.pbr.metallic = 0.0f;
    s.mesh.material.pbr.specular = 0.0f;
    check("principled matte == POV diffuse", umbreon::render(s, o).color == pov.color);
    s.mesh.material.pbr.metallic = 1.0f;
    s.mesh.material.pbr.specular = 0.5f;
    check("principled metal differs", umbreon::render(s, o).color != pov.color);
  }
  {
    // object-space edges through RenderOptions (test_object_space_edges.cpp:
    // 566-700): they change the image, the caller's scene is untouched, and
    // enabling both edge methods throws
    umbreon::Scene s;
    s.camera = orthoCam();
    s.background = {1, 1, 1};
    umbreon::Sphere sp;
    sp.center = {0, 0, 0};
    sp.radius = 1.0f;
    sp.color = {0.3f, 0.4f, 0.5f, 1.0f};
    s.spheres.push_back(sp);
    umbreon::RenderOptions r;
    r.width = 32;
    r.height = 32;
    const umbreon::FrameResult plain = umbreon::render(s, r);
    r.objectSpaceEdges.enable = true;
    r.objectSpaceEdges.width = 0.2f;
    r.objectSpaceEdges.raise = 0.15f;
    const umbreon::FrameResult edged = umbreon::render(s, r);
    check("object-space edges change the image", edged.color != plain.color);
    check("object-space edges leave the scene untouched", s.cylinders.empty());
    r.strokeEdges.enable = true;
    bool threw = false;
    try {
      umbreon::render(s, r);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    check("both edge methods throw", threw);
  }
  std::printf("%d failure(s)\n", failures);
  return failures == 0 ? 0 : 1;
}
