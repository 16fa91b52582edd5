// SPDX-License-Identifier: Apache-2.0
// Deterministic synthetic LP generators (see generate.cpp for the constructions).
#pragma once

#include <cstdint>

#include "sankhya/model.hpp"

namespace sankhya::hip_engine {

struct GenSpec {
  std::int32_t kind = 0;     // 0 uniform-equality, 1 zipf, 2 blend
  std::int32_t m = 0, n = 0;
  std::int32_t per_col = 5;  // distinct random rows per column
  double param = 1.5;        // Zipf exponent (kind 1)
  std::uint64_t seed = 1;
};

[[nodiscard]] Model generate(const GenSpec& spec);

}  // namespace sankhya::hip_engine
