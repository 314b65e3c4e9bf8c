// Host-side support of the public API: log sink, progress accounting, and the
// image / fog helpers callers may apply to frames they hold
// (log.cpp:37-64, render_progress.cpp, image_ops.cpp:9-79 and fog.cpp:7-36 of
// the reference define their behaviour).

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "log.hpp"
#include "postprocess/fog.hpp"
#include "postprocess/image_ops.hpp"
#include "render/render_progress.hpp"

namespace umbreon {

// ---------------------------------------------------------------- log
namespace {
std::mutex& logMutex() {
  static std::mutex m;
  return m;
}
LogSink& sinkRef() {
  static LogSink s;
  return s;
}
}  // namespace

void setLogSink(LogSink sink) {
  std::lock_guard<std::mutex> lk(logMutex());
  sinkRef() = std::move(sink);
}

void logMessage(LogLevel level, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(logMutex());
  if (sinkRef()) {
    sinkRef()(level, buf);
  } else if (level == LogLevel::Warning) {
    std::fprintf(stderr, "umbreon: warning: %s\n", buf);
  }
}

// ---------------------------------------------------------------- progress
namespace {
int slotOf(RenderPhase p) {
  switch (p) {
    case RenderPhase::Setup: return 0;
    case RenderPhase::CoarseAo: return 1;
    case RenderPhase::Primary: return 2;
    case RenderPhase::GlobalIllum: return 3;
    case RenderPhase::Edges: return 4;
    case RenderPhase::Postprocess: return 5;
    default: return -1;
  }
}
}  // namespace

float RenderProgress::fraction() const noexcept {
  const int ph = phase_.load(std::memory_order_relaxed);
  if (ph == static_cast<int>(RenderPhase::Done)) return 1.0f;
  uint32_t share[kSlots];
  uint64_t sum = 0;
  for (int i = 0; i < kSlots; ++i) {
    share[i] = share_[i].load(std::memory_order_relaxed);
    sum += share[i];
  }
  if (sum == 0) return 0.0f;
  double within = 0.0;
  const int slot = slotOf(static_cast<RenderPhase>(ph));
  if (slot >= 0) {
    for (int i = 0; i < slot; ++i) within += share[i];
    const uint64_t tot = std::max<uint64_t>(1, total_.load(std::memory_order_relaxed));
    const uint64_t dn = std::min(tot, done_.load(std::memory_order_relaxed));
    within += share[slot] * (static_cast<double>(dn) / static_cast<double>(tot));
  }
  within /= static_cast<double>(sum);
  const uint64_t passes = std::max<uint64_t>(1, passes_.load(std::memory_order_relaxed));
  const uint64_t pass = std::min(passes - 1, pass_.load(std::memory_order_relaxed));
  const double f = (static_cast<double>(pass) + within) / static_cast<double>(passes);
  return static_cast<float>(std::min(1.0, std::max(0.0, f)));
}

RenderPhasePlan RenderProgress::phasePlan() const noexcept {
  RenderPhasePlan p;
  float* f[kSlots] = {&p.setup, &p.coarseAo, &p.primary, &p.globalIllum, &p.edges, &p.postprocess};
  for (int i = 0; i < kSlots; ++i) *f[i] = share_[i].load(std::memory_order_relaxed) * 1e-3f;
  return p;
}

void RenderProgress::setPhasePlan(const RenderPhasePlan& plan) noexcept {
  const float in[kSlots] = {plan.setup, plan.coarseAo, plan.primary, plan.globalIllum, plan.edges,
                            plan.postprocess};
  double sum = 0.0;
  for (float v : in)
    if (std::isfinite(v) && v > 0.0f) sum += v;
  if (!(sum > 0.0)) return;  // no information: keep the current plan
  for (int i = 0; i < kSlots; ++i) {
    const double v = (std::isfinite(in[i]) && in[i] > 0.0f) ? in[i] : 0.0;
    share_[i].store(static_cast<uint32_t>(v / sum * 1000.0 + 0.5), std::memory_order_relaxed);
  }
}

void RenderProgress::beginPass(std::uint64_t passIndex, std::uint64_t passCount) noexcept {
  passes_.store(passCount ? passCount : 1, std::memory_order_relaxed);
  pass_.store(passIndex, std::memory_order_relaxed);
  total_.store(1, std::memory_order_relaxed);
  done_.store(0, std::memory_order_relaxed);
  phase_.store(static_cast<int>(RenderPhase::Idle), std::memory_order_relaxed);
}

void RenderProgress::beginPhase(RenderPhase p, std::uint64_t unitsTotal) noexcept {
  total_.store(unitsTotal ? unitsTotal : 1, std::memory_order_relaxed);
  done_.store(0, std::memory_order_relaxed);
  phase_.store(static_cast<int>(p), std::memory_order_relaxed);
}

void RenderProgress::advanceTo(std::uint64_t unitsDone) noexcept {
  std::uint64_t cur = done_.load(std::memory_order_relaxed);
  while (cur < unitsDone && !done_.compare_exchange_weak(cur, unitsDone, std::memory_order_relaxed)) {
  }
}

void RenderProgress::markDone() noexcept {
  pass_.store(0, std::memory_order_relaxed);
  passes_.store(1, std::memory_order_relaxed);
  total_.store(1, std::memory_order_relaxed);
  done_.store(1, std::memory_order_relaxed);
  phase_.store(static_cast<int>(RenderPhase::Done), std::memory_order_relaxed);
}

// Display names of the phases, indexed by the enum value; the strings are
// the reference's (src/umbreon/render/render_progress.cpp:109-121), which a
// UI may show verbatim.
const char* toString(RenderPhase p) noexcept {
  static const char* const kNames[] = {"Idle", "Setup", "CoarseAO", "Primary", "GI", "Edges", "Postprocess", "Done"};
  const int i = static_cast<int>(p);
  return i >= 0 && i < static_cast<int>(sizeof(kNames) / sizeof(kNames[0])) ? kNames[i] : "Unknown";
}

// ---------------------------------------------------------------- image ops
std::vector<float> boxDownsample(const std::vector<float>& src, int w, int h, int channels, int ss) {
  const int ow = w / ss, oh = h / ss;
  std::vector<float> out(static_cast<std::size_t>(ow) * oh * channels, 0.0f);
  const float inv = 1.0f / static_cast<float>(ss * ss);
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x)
      for (int c = 0; c < channels; ++c) {
        float acc = 0.0f;
        for (int dy = 0; dy < ss; ++dy)
          for (int dx = 0; dx < ss; ++dx) {
            const float v = src[(static_cast<std::size_t>(y * ss + dy) * w + (x * ss + dx)) * channels + c];
            acc += std::isfinite(v) ? v : 0.0f;
          }
        out[(static_cast<std::size_t>(y) * ow + x) * channels + c] = acc * inv;
      }
  return out;
}

void applyAssumedGamma(FrameResult& frame, float g) {
  if (std::fabs(g - 1.0f) <= 1.0e-4f) return;
  const std::size_t n = static_cast<std::size_t>(frame.width) * frame.height;
  for (std::size_t i = 0; i < n && 4 * i + 3 < frame.color.size(); ++i)
    for (int c = 0; c < 3; ++c) {
      float& v = frame.color[4 * i + c];
      v = v > 0.0f ? std::pow(v, g) : 0.0f;
    }
}

float srgbEncodeF(float v) {
  v = std::min(1.0f, std::max(0.0f, v));
  return v <= 0.0031308f ? 12.92f * v : 1.055f * std::pow(v, 1.0f / 2.4f) - 0.055f;
}

float srgbDecodeF(float s) {
  s = std::min(1.0f, std::max(0.0f, s));
  return s <= 0.04045f ? s / 12.92f : std::pow((s + 0.055f) / 1.055f, 2.4f);
}

std::vector<std::uint8_t> srgbEncode8(const FrameResult& frame, int channels) {
  const std::size_t n = static_cast<std::size_t>(frame.width) * frame.height;
  std::vector<std::uint8_t> out(n * channels);
  auto q = [](float v) {
    const int i = static_cast<int>(v * 255.0f + 0.5f);
    return static_cast<std::uint8_t>(std::min(255, std::max(0, i)));
  };
  for (std::size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3 && c < channels; ++c) out[i * channels + c] = q(srgbEncodeF(frame.color[4 * i + c]));
    if (channels == 4) out[4 * i + 3] = q(std::min(1.0f, std::max(0.0f, frame.color[4 * i + 3])));
  }
  return out;
}

// ---------------------------------------------------------------- fog
float fogFactor(const Fog& fog, float z) {
  if (!fog.enabled) return 1.0f;
  const float denom = fog.end - fog.start;
  if (denom <= 0.0f) return z < fog.end ? 1.0f : 0.0f;
  const float f = (fog.end - z) / denom;
  return f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
}

void applyFog(const Fog& fog, int width, int height, int channels, float* color, const float* viewZ,
              bool transparentBackground) {
  if (!fog.enabled || !color || !viewZ) return;
  const std::size_t n = static_cast<std::size_t>(width) * height;
  for (std::size_t i = 0; i < n; ++i) {
    const float z = viewZ[i];
    if (z <= 0.0f) continue;
    const float f = fogFactor(fog, z);
    float* c = color + i * channels;
    if (transparentBackground) {
      if (channels == 4) c[3] *= f;
    } else {
      c[0] = fog.color.x * (1.0f - f) + c[0] * f;
      c[1] = fog.color.y * (1.0f - f) + c[1] * f;
      c[2] = fog.color.z * (1.0f - f) + c[2] * f;
    }
  }
}

}  // namespace umbreon
