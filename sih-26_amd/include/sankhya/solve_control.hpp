// SPDX-License-Identifier: Apache-2.0
// Progress callback and interruption (reference: include/sankhya/solve_control.hpp:1-107).
#pragma once

#include <atomic>
#include <chrono>
#include <functional>

#include "sankhya/types.hpp"

namespace sankhya {

struct Progress {
  enum class Phase { kPresolve, kLp, kMip, kPolish };
  Phase phase = Phase::kLp;
  Count iterations = 0;
  double objective = 0.0;
  double best_bound = 0.0;
  double elapsed_seconds = 0.0;
};

/// Return non-zero to interrupt the solve.
using ProgressCallback = std::function<int(const Progress&)>;

class SolveControl {
 public:
  static constexpr std::chrono::milliseconds kCallbackInterval{100};
  ProgressCallback progress_callback;

  [[nodiscard]] bool callback_due() noexcept {
    const auto now = std::chrono::steady_clock::now();
    if (callback_seen_ && now - last_callback_ < kCallbackInterval) return false;
    callback_seen_ = true;
    last_callback_ = now;
    return true;
  }
  void interrupt() noexcept { interrupt_requested_.store(true, std::memory_order_relaxed); }
  [[nodiscard]] bool interruption_requested() const noexcept {
    return interrupt_requested_.load(std::memory_order_relaxed);
  }
  void reset() noexcept {
    interrupt_requested_.store(false, std::memory_order_relaxed);
    callback_seen_ = false;
  }

 private:
  std::atomic<bool> interrupt_requested_{false};
  bool callback_seen_ = false;
  std::chrono::steady_clock::time_point last_callback_{};
};

}  // namespace sankhya
