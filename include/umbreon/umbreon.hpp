// umbreon public entry points, MI355X backend.
//
// Same API as libumbreon (src/umbreon/umbreon.hpp): build a Scene, call
// render(), get a linear HDR FrameResult. Behind it the per-ray path runs on
// one gfx950 GPU through the C-ABI in umbreon_gpu.h (libumbreon_gpu.so): the
// scene is flattened to structure-of-arrays, uploaded with its BVH, traced by
// the wavefront kernels and resolved (fog, supersample, gamma) on the device.
#pragma once

#include <chrono>
#include <memory>

#include "render/render_types.hpp"
#include "scene.hpp"

#define UMBREON_VERSION_MAJOR 0
#define UMBREON_VERSION_MINOR 1
#define UMBREON_VERSION_PATCH 0

namespace umbreon {

// Synchronous render on the calling thread. Throws std::runtime_error on a
// device failure or an option the GPU path does not implement.
FrameResult render(const Scene& scene, const RenderOptions& opt);

// As above with progress reporting and cooperative cancellation: a cancelled
// render returns a partial FrameResult with cancelled == true.
FrameResult render(const Scene& scene, const RenderOptions& opt, RenderProgress& progress);

// Handle of a render running on a background thread.
class RenderTask {
 public:
  RenderTask(RenderTask&&) noexcept;
  RenderTask& operator=(RenderTask&&) noexcept;
  ~RenderTask();  // requests cancel and joins
  RenderTask(const RenderTask&) = delete;
  RenderTask& operator=(const RenderTask&) = delete;

  float progress() const noexcept;
  RenderPhase phase() const noexcept;
  bool done() const noexcept;
  void cancel() noexcept;
  bool wait_for(std::chrono::milliseconds timeout) const;
  void wait() const;
  // Join and return the frame; rethrows the render's exception. Call once.
  FrameResult get();

 private:
  struct Impl;
  explicit RenderTask(std::unique_ptr<Impl> impl) noexcept;
  friend RenderTask renderAsync(Scene scene, RenderOptions opt);
  std::unique_ptr<Impl> p_;
};

RenderTask renderAsync(Scene scene, RenderOptions opt);

}  // namespace umbreon
