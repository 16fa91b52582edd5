// SPDX-License-Identifier: Apache-2.0
// The solve() seam for LPs (reference: src/core/solve.cpp:500-561 and the LP branch
// :695-935), reduced to what this engine needs: log options, the LP check, engine selection
// (PDHG on the MI355X is the only engine), and a guard that turns an exception into a status.
#include <exception>
#include <new>

#include "engine_stats.hpp"
#include "sankhya/pdhg.hpp"

namespace sankhya {


Solution solve(const Model& model, const Options& options, SolveControl* control) {
  Logger logger(options.get_bool("log_to_console") ? stdout : nullptr,
                log_level_of(options.get_string("log_level")));
  if (!options.get_string("progress_out").empty())
    logger.enable_progress_output(options.get_string("progress_out"));
  Solution solution;
  if (model.has_integrality() || model.has_quadratic_objective()) {
    solution.status = SolveStatus::kModelError;
    solution.message = "this engine solves linear programs only (no integrality, no quadratic objective)";
    return solution;
  }
  try {
    solution = hip_engine::run(model, options, logger, control);
  } catch (const std::bad_alloc&) {
    solution = Solution{};
    solution.status = SolveStatus::kNumericalError;
    solution.algorithm = "pdhg-gpu";
    solution.message = "ran out of memory inside PDHG";
  } catch (const std::exception& e) {
    solution = Solution{};
    solution.status = SolveStatus::kNumericalError;
    solution.algorithm = "pdhg-gpu";
    solution.message = std::string("PDHG failed: ") + e.what();
  }
  return solution;
}

}  // namespace sankhya
