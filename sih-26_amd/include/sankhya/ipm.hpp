// SPDX-License-Identifier: Apache-2.0
// The interior-point warm start (reference: include/sankhya/ipm.hpp:25-29) and the hook through
// which solve() reaches an interior-point engine for the polish of a PDHG answer. In SANKHYA
// the dispatcher calls ipm::solve_ipm itself (src/core/solve.cpp:239-308, :784-810). This
// library has no interior-point engine, so inside SANKHYA the dispatcher keeps its own polish
// and this hook stays empty; a standalone caller that links an interior-point engine installs
// it here and solve() then runs the same handoff.
#pragma once

#include <vector>

#include "sankhya/logging.hpp"
#include "sankhya/model.hpp"
#include "sankhya/options.hpp"
#include "sankhya/solve_control.hpp"
#include "sankhya/timer.hpp"

namespace sankhya::ipm {

/// A starting point in the ORIGINAL model's units and sign convention (the ones a Solution
/// carries): PDHG's point, row duals and reduced costs.
struct WarmStart {
  std::vector<double> col_value;
  std::vector<double> row_dual;
  std::vector<double> col_dual;
};

}  // namespace sankhya::ipm

namespace sankhya {

/// The signature of ipm::solve_ipm with a SolveControl and a warm start (reference
/// include/sankhya/ipm.hpp:38-40).
using PolishEngine = Solution (*)(const Model& model, const Options& options, Logger& logger,
                                  SolveControl* control, const ipm::WarmStart* warm);

/// Installs the engine solve() polishes with, process-wide; nullptr removes it.
void install_polish_engine(PolishEngine engine) noexcept;
[[nodiscard]] PolishEngine installed_polish_engine() noexcept;

namespace detail {

/// PDHG's share of a finite time limit when a polish is to follow (reference solve.cpp:311).
constexpr double kPdhgShareOfTheTimeLimit = 0.7;

/// The polish step of solve() (reference solve.cpp:239-308): when pdhg_polish is on, an engine
/// is installed and `first` is not optimal but holds a point of the model's size, the point
/// goes to the engine as its warm start with polish_iteration_limit iterations and the smaller
/// of polish_max_seconds and what time_limit has left (polish_max_seconds alone in
/// deterministic mode). The polished answer replaces `first` only when it is optimal, or
/// feasible with smaller scaled violations; its iterations are then the sum of both phases,
/// polish_iterations the second's, and its algorithm `first`'s with "+ipm".
void polish_with_the_interior_point(Solution* first, const Model& model, const Options& options,
                                    Logger& logger, SolveControl* control, const Timer& timer);

}  // namespace detail
}  // namespace sankhya
