// SPDX-License-Identifier: Apache-2.0
// Wall-clock timer (reference: include/sankhya/timer.hpp).
#pragma once

#include <chrono>

namespace sankhya {

class Timer {
 public:
  Timer() : start_(std::chrono::steady_clock::now()) {}
  void reset() { start_ = std::chrono::steady_clock::now(); }
  [[nodiscard]] double elapsed_seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
  }

 private:
  std::chrono::steady_clock::time_point start_;
};

}  // namespace sankhya
