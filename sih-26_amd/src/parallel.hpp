// SPDX-License-Identifier: Apache-2.0
// Host helper: split [0, count) into `threads` contiguous ranges, one per thread.
#pragma once

#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace sankhya::hip_engine {

// f(begin, end) on each range; range k is [k * chunk, (k + 1) * chunk) with
// chunk = ceil(count / threads), so a caller can index per-thread state by begin / chunk. The
// calling thread runs the first range.
template <class F>
void parallel_ranges(std::size_t count, int threads, F&& f) {
  if (threads <= 1 || count < 2) {
    f(std::size_t{0}, count);
    return;
  }
  const std::size_t t = static_cast<std::size_t>(threads);
  const std::size_t chunk = (count + t - 1) / t;
  std::vector<std::thread> pool;
  for (std::size_t k = 1; k < t; ++k) {
    const std::size_t b = k * chunk, e = std::min(count, b + chunk);
    if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
  }
  f(std::size_t{0}, std::min(count, chunk));
  for (std::thread& th : pool) th.join();
}

}  // namespace sankhya::hip_engine
