// SPDX-License-Identifier: Apache-2.0
// Core scalar and index types of the SANKHYA API, as the PDHG path uses them
// (reference: include/sankhya/types.hpp:14-116).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>

namespace sankhya {

/// Row, column and nonzero index (reference types.hpp:17).
using Index = std::int32_t;
/// Iteration and other unbounded counters (reference types.hpp:20).
using Count = std::int64_t;

inline constexpr Index kMaxNonzeros = std::numeric_limits<Index>::max();

[[nodiscard]] inline bool nonzero_count_fits(std::size_t count) noexcept {
  return count <= static_cast<std::size_t>(kMaxNonzeros);
}

/// An absent bound (reference types.hpp:82).
inline constexpr double kInfinity = std::numeric_limits<double>::infinity();
/// MPS files write 1e30 for infinity; anything at or beyond it is an absent bound.
inline constexpr double kMpsInfinity = 1e30;

[[nodiscard]] inline bool is_infinite(double v) noexcept {
  return std::isinf(v) || std::fabs(v) >= kMpsInfinity;
}
[[nodiscard]] inline bool is_finite_bound(double v) noexcept { return !is_infinite(v); }
[[nodiscard]] inline double normalize_infinity(double v) noexcept {
  if (!is_infinite(v)) return v;
  return v > 0.0 ? kInfinity : -kInfinity;
}
[[nodiscard]] inline double normalize_zero(double v) noexcept { return v == 0.0 ? 0.0 : v; }

}  // namespace sankhya
