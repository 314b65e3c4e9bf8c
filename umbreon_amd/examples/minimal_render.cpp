// Minimal consumer of the umbreon C++ API on the MI355X backend: one lit quad,
// AO and shadows on, 2x supersampling, written as a binary PPM. Mirrors the
// reference's examples/minimal_render.cpp usage (same public calls).
#include <cstdio>
#include <vector>

#include "postprocess/image_ops.hpp"
#include "umbreon.hpp"

int main() {
  umbreon::Scene scene;
  umbreon::Mesh& mesh = scene.mesh;
  const umbreon::Vec3 corners[6] = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
  for (const umbreon::Vec3& c : corners) {
    mesh.positions.push_back(c);
    mesh.normals.push_back({0.0f, 0.0f, 1.0f});
    mesh.colors.push_back({0.2f, 0.4f, 0.8f, 1.0f});
  }
  mesh.material.ambient = 0.2f;
  mesh.material.diffuse = 0.8f;
  scene.camera.position = {0.0f, 0.0f, 5.0f};
  scene.camera.direction = {0.0f, 0.0f, -1.0f};
  scene.camera.up = {0.0f, 1.0f, 0.0f};
  scene.camera.orthographic = true;
  scene.camera.height = 2.5f;
  umbreon::DistantLight light;
  light.direction = {0.0f, 0.0f, -1.0f};
  light.intensity = 1.0f;
  scene.lights.push_back(light);

  umbreon::RenderOptions opt;
  opt.width = 256;
  opt.height = 256;
  opt.supersample = 2;
  opt.aoSamples = 16;
  opt.shadows = true;

  umbreon::FrameResult frame = umbreon::render(scene, opt);
  std::printf("rendered %dx%d in %.3f s (%zu effective triangles)\n", frame.width, frame.height,
              frame.renderSeconds, frame.effectiveTriangles);
  const std::size_t c = (static_cast<std::size_t>(frame.height / 2) * frame.width + frame.width / 2) * 4;
  std::printf("center pixel %.4f %.4f %.4f %.4f\n", frame.color[c], frame.color[c + 1], frame.color[c + 2],
              frame.color[c + 3]);
  std::vector<std::uint8_t> rgb = umbreon::srgbEncode8(frame, 3);
  if (FILE* f = std::fopen("minimal.ppm", "wb")) {
    std::fprintf(f, "P6\n%d %d\n255\n", frame.width, frame.height);
    std::fwrite(rgb.data(), 1, rgb.size(), f);
    std::fclose(f);
  }
  return 0;
}
