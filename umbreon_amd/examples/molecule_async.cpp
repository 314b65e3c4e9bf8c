// A CueMol-style consumer of the umbreon C++ API on the MI355X backend: a
// ball-and-stick benzene molecule (12 atom spheres, 12 bonds as pairs of
// capped half-bond cylinders coloured by their atom) rendered in the
// background with renderAsync while the caller polls the phase and progress,
// then written as a binary PPM ("benzene.ppm").
//
// Exit status 0 when the frame is complete, the ring's centre shows the
// background through the hole and the carbon atoms are lit.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "postprocess/image_ops.hpp"
#include "umbreon.hpp"

namespace {

struct Atom {
  umbreon::Vec3 p;
  umbreon::Vec4 rgba;
  float radius;
};

// Planar hexagon of carbons (C-C 1.39 A) with hydrogens 1.09 A further out.
std::vector<Atom> benzene() {
  std::vector<Atom> atoms;
  const float kPi = 3.14159265f;
  for (int i = 0; i < 6; ++i) {
    const float a = kPi / 3.0f * static_cast<float>(i);
    const float c = std::cos(a), s = std::sin(a);
    atoms.push_back({{1.39f * c, 1.39f * s, 0.0f}, {0.35f, 0.35f, 0.35f, 1.0f}, 0.40f});
    atoms.push_back({{2.48f * c, 2.48f * s, 0.0f}, {0.95f, 0.95f, 0.95f, 1.0f}, 0.25f});
  }
  return atoms;
}

umbreon::Scene buildScene() {
  umbreon::Scene scene;
  const std::vector<Atom> atoms = benzene();
  umbreon::Material finish;
  finish.diffuse = 0.7f;
  finish.specular = 0.3f;
  finish.roughness = 0.01f;
  for (const Atom& a : atoms) {
    umbreon::Sphere s;
    s.center = a.p;
    s.radius = a.radius;
    s.color = a.rgba;
    s.material = finish;
    scene.spheres.push_back(s);
  }
  // bonds: ring C(i)-C(i+1) and C(i)-H(i); each drawn as two half cylinders
  auto bond = [&](int i, int j) {
    const umbreon::Vec3 m{0.5f * (atoms[i].p.x + atoms[j].p.x), 0.5f * (atoms[i].p.y + atoms[j].p.y),
                          0.5f * (atoms[i].p.z + atoms[j].p.z)};
    for (int k : {i, j}) {
      umbreon::Cylinder c;
      c.p0 = atoms[k].p;
      c.p1 = m;
      c.radius = 0.15f;
      c.color = atoms[k].rgba;
      c.material = finish;
      c.open = false;
      scene.cylinders.push_back(c);
    }
  };
  for (int i = 0; i < 6; ++i) {
    bond(2 * i, 2 * ((i + 1) % 6));
    bond(2 * i, 2 * i + 1);
  }
  scene.camera.position = {0.0f, -3.0f, 9.0f};
  scene.camera.direction = {0.0f, 3.0f, -9.0f};
  scene.camera.up = {0.0f, 1.0f, 0.0f};
  scene.camera.fovy = 40.0f;
  umbreon::DistantLight key;
  key.direction = {-0.4f, -0.5f, -1.0f};
  key.intensity = 0.9f;
  scene.lights.push_back(key);
  scene.background = {0.05f, 0.07f, 0.12f};
  return scene;
}

}  // namespace

int main() {
  umbreon::RenderOptions opt;
  opt.width = 320;
  opt.height = 240;
  opt.supersample = 2;
  opt.aoSamples = 8;
  opt.aoDistance = 3.0f;
  opt.shadows = true;

  umbreon::RenderTask task = umbreon::renderAsync(buildScene(), opt);
  std::string lastPhase;
  while (!task.wait_for(std::chrono::milliseconds(2))) {
    const std::string ph = umbreon::toString(task.phase());
    if (ph != lastPhase) {
      std::printf("phase %-12s %5.1f%%\n", ph.c_str(), 100.0 * task.progress());
      lastPhase = ph;
    }
  }
  umbreon::FrameResult frame = task.get();
  std::printf("frame %dx%d, %.3f ms of trace, cancelled %d\n", frame.width, frame.height,
              1e3 * frame.renderSeconds, frame.cancelled ? 1 : 0);

  std::vector<std::uint8_t> rgb = umbreon::srgbEncode8(frame, 3);
  if (FILE* f = std::fopen("benzene.ppm", "wb")) {
    std::fprintf(f, "P6\n%d %d\n255\n", frame.width, frame.height);
    std::fwrite(rgb.data(), 1, rgb.size(), f);
    std::fclose(f);
  }

  // The camera looks at the ring's centre through its hole; the first carbon
  // (+x) is lit by the key light.
  auto px = [&](int x, int y) { return &frame.color[(static_cast<std::size_t>(y) * frame.width + x) * 4]; };
  const float* centre = px(frame.width / 2, frame.height / 2);
  const bool hole = std::fabs(centre[0] - 0.05f) < 1e-3f && std::fabs(centre[2] - 0.12f) < 1e-3f;
  float brightest = 0.0f;
  for (int y = 0; y < frame.height; ++y)
    for (int x = 0; x < frame.width; ++x) brightest = std::fmax(brightest, px(x, y)[0]);
  std::printf("centre %.4f %.4f %.4f (background through the ring: %s), brightest red %.3f\n", centre[0],
              centre[1], centre[2], hole ? "yes" : "no", brightest);
  return (!frame.cancelled && hole && brightest > 0.3f) ? 0 : 1;
}
