// SPDX-License-Identifier: Apache-2.0
// The tolerances the PDHG path reads (reference: include/sankhya/tolerances.hpp:19-182).
#pragma once

#include "sankhya/types.hpp"

namespace sankhya::tol {

/// Largest row or column bound violation of a feasible point (reference tolerances.hpp:19).
inline constexpr double kPrimalFeasibility = 1e-7;
/// Largest reduced-cost sign violation of a dual-feasible point (tolerances.hpp:22).
inline constexpr double kDualFeasibility = 1e-7;
/// Relative duality gap an optimality claim may carry (tolerances.hpp:108).
inline constexpr double kDualityGap = 1e-9;
/// Largest |multiplier| * slack an optimality claim may carry (tolerances.hpp:118).
inline constexpr double kComplementarity = 1e-6;
/// Entries below this are dropped when a matrix is finalised (tolerances.hpp:126).
inline constexpr double kZeroDrop = 1e-11;
/// The two PDHG reporting tolerances (tolerances.hpp:181-182).
inline constexpr double kPdhgLoose = 1e-4;
inline constexpr double kPdhgTight = 1e-8;
/// Integrality, used by the quality report on integer columns (tolerances.hpp:25).
inline constexpr double kIntegrality = 1e-6;
/// MIP gap targets; a solution file records them in its header (tolerances.hpp:28,31).
inline constexpr double kMipRelativeGap = 1e-4;
inline constexpr double kMipAbsoluteGap = 1e-6;

}  // namespace sankhya::tol
