// Lock-free render progress + cooperative cancel (source-compatible with
// libumbreon's render/render_progress.hpp). One thread renders, others poll
// fraction()/phase() or call requestCancel(); the GPU backend checks the
// cancel flag between wavefront passes.
#pragma once

#include <atomic>
#include <cstdint>

namespace umbreon {

enum class RenderPhase : int { Idle = 0, Setup, CoarseAo, Primary, GlobalIllum, Edges, Postprocess, Done };

// Expected share of render time per phase (normalized when stored).
struct RenderPhasePlan {
  float setup = 0.05f;
  float coarseAo = 0.05f;
  float primary = 0.65f;
  float globalIllum = 0.15f;
  float edges = 0.05f;
  float postprocess = 0.05f;
};

namespace detail {
class ProgressAccess;
}

class RenderProgress {
 public:
  RenderProgress() = default;
  RenderProgress(const RenderProgress&) = delete;
  RenderProgress& operator=(const RenderProgress&) = delete;

  // Overall completion in [0, 1]: finished passes, then within the current
  // pass the plan shares of finished phases plus the current phase's share
  // scaled by unitsDone / unitsTotal.
  float fraction() const noexcept;
  RenderPhasePlan phasePlan() const noexcept;
  RenderPhase phase() const noexcept { return static_cast<RenderPhase>(phase_.load(std::memory_order_relaxed)); }
  std::uint64_t unitsDone() const noexcept { return done_.load(std::memory_order_relaxed); }
  std::uint64_t unitsTotal() const noexcept { return total_.load(std::memory_order_relaxed); }
  void requestCancel() noexcept { cancel_.store(1, std::memory_order_relaxed); }
  bool cancelRequested() const noexcept { return cancel_.load(std::memory_order_relaxed) != 0; }
  void setPhasePlan(const RenderPhasePlan& plan) noexcept;
  void beginPass(std::uint64_t passIndex, std::uint64_t passCount) noexcept;
  void beginPhase(RenderPhase p, std::uint64_t unitsTotal = 1) noexcept;
  void advance(std::uint64_t n) noexcept { done_.fetch_add(n, std::memory_order_relaxed); }
  void advanceTo(std::uint64_t unitsDone) noexcept;
  void markDone() noexcept;

 private:
  friend class detail::ProgressAccess;
  static constexpr int kSlots = 6;
  std::atomic<int> phase_{static_cast<int>(RenderPhase::Idle)};
  std::atomic<std::uint64_t> done_{0};
  std::atomic<std::uint64_t> total_{1};
  std::atomic<std::uint64_t> pass_{0};
  std::atomic<std::uint64_t> passes_{1};
  std::atomic<std::uint32_t> share_[kSlots] = {{50}, {50}, {650}, {150}, {50}, {50}};  // permille
  std::atomic<std::int32_t> cancel_{0};
};

const char* toString(RenderPhase p) noexcept;

}  // namespace umbreon
