// SPDX-License-Identifier: Apache-2.0
// The solve() seam for LPs (reference: src/core/solve.cpp:500-561 and the LP branch
// :695-935), reduced to what this engine needs: log options, the LP check, engine selection
// (PDHG on the MI355X is the only engine), a guard that turns an exception into a status, and
// the interior-point polish of a PDHG answer through an installed engine (:239-308, :784-810).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <new>
#include <string>

#include "engine_stats.hpp"
#include "sankhya/ipm.hpp"
#include "sankhya/pdhg.hpp"

namespace sankhya {

namespace {

std::atomic<PolishEngine> g_polish{nullptr};

double worst_scaled(const Solution& s) {
  return std::max(s.primal_infeasibility_scaled, s.dual_infeasibility_scaled);
}

}  // namespace

void install_polish_engine(PolishEngine engine) noexcept { g_polish.store(engine); }

PolishEngine installed_polish_engine() noexcept { return g_polish.load(); }

namespace detail {

void polish_with_the_interior_point(Solution* first, const Model& model, const Options& options,
                                    Logger& logger, SolveControl* control, const Timer& timer) {
  const PolishEngine engine = installed_polish_engine();
  if (engine == nullptr || !options.get_bool("pdhg_polish")) return;
  if (first->status == SolveStatus::kOptimal || !claims_a_point(first->status)) return;
  const auto n = static_cast<std::size_t>(model.num_cols());
  const auto m = static_cast<std::size_t>(model.num_rows());
  if (first->col_value.size() != n || first->row_dual.size() != m || first->col_dual.size() != n)
    return;

  // The polish spends budgets of its own: polish_iteration_limit iterations, and the smaller
  // of polish_max_seconds and what the time limit has left (the first alone when
  // deterministic, where the clock may not decide anything).
  Options polish = options;
  polish.set_int("iteration_limit", options.get_int("polish_iteration_limit"));
  double budget = options.get_double("polish_max_seconds");
  const double time_limit = options.get_double("time_limit");
  if (!options.get_bool("deterministic") && std::isfinite(time_limit)) {
    const double remaining = time_limit - timer.elapsed_seconds();
    if (remaining <= 0.0) {
      first->message += "; no time left for the interior-point polish";
      return;
    }
    budget = std::min(budget, remaining);
  }
  polish.set_double("time_limit", budget);
  logger.info("Polish: handing PDHG's point to the interior point, {} iterations at most",
              polish.get_int("iteration_limit"));
  const ipm::WarmStart warm{first->col_value, first->row_dual, first->col_dual};
  Solution polished = engine(model, polish, logger, control, &warm);

  const bool better = polished.status == SolveStatus::kOptimal ||
                      (polished.status == SolveStatus::kFeasible &&
                       worst_scaled(polished) < worst_scaled(*first));
  if (!better) {
    first->polish_iterations = polished.iterations;
    first->message += std::string("; the interior-point polish did not improve it (") +
                      to_string(polished.status) + " after " +
                      std::to_string(polished.iterations) + " iterations: " + polished.message +
                      ")";
    return;
  }
  polished.polish_iterations = polished.iterations;
  polished.iterations += first->iterations;
  polished.algorithm = first->algorithm + "+ipm";
  polished.message = first->message + "; polished by the interior point in " +
                     std::to_string(polished.polish_iterations) + " iterations" +
                     (polished.message.empty() ? std::string() : " (" + polished.message + ")");
  *first = std::move(polished);
}

}  // namespace detail

Solution solve(const Model& model, const Options& options, SolveControl* control) {
  const Timer timer;
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
  // With a polish to follow, PDHG gets its share of a finite time limit and the polish the
  // rest (reference solve.cpp:784-791).
  Options first_pass = options;
  const double time_limit = options.get_double("time_limit");
  if (installed_polish_engine() != nullptr && options.get_bool("pdhg_polish") &&
      !options.get_bool("deterministic") && time_limit > 0.0 && std::isfinite(time_limit)) {
    first_pass.set_double("time_limit", time_limit * detail::kPdhgShareOfTheTimeLimit);
  }
  try {
    solution = hip_engine::run(model, first_pass, logger, control);
    detail::polish_with_the_interior_point(&solution, model, options, logger, control, timer);
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
