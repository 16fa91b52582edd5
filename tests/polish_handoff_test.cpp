// The interior-point polish handoff of solve() (the reference's polish_with_the_interior_point,
// src/core/solve.cpp:239-308), run through sankhya::detail::polish_with_the_interior_point with
// a stub in place of the interior-point engine. Built and run by tests/test_host.py; prints
// one line per failed check and exits non-zero when any fails.
#include <cmath>
#include <cstdio>
#include <string>

#include "sankhya/ipm.hpp"

using namespace sankhya;

namespace {

int failures = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

int calls = 0;
ipm::WarmStart seen;
std::int64_t seen_iterations = -1;
double seen_time = -1.0;
Solution reply;

Solution stub(const Model&, const Options& options, Logger&, SolveControl*, const ipm::WarmStart* warm) {
  ++calls;
  seen = *warm;
  seen_iterations = options.get_int("iteration_limit");
  seen_time = options.get_double("time_limit");
  return reply;
}

Model two_by_one() {
  Model m;
  m.col_cost = {1.0, 2.0};
  m.col_lower = {0.0, 0.0};
  m.col_upper = {1.0, 1.0};
  m.row_lower = {1.0};
  m.row_upper = {1.0};
  return m;
}

Solution pdhg_answer(SolveStatus status) {
  Solution s;
  s.status = status;
  s.col_value = {0.25, 0.75};
  s.row_dual = {-1.5};
  s.col_dual = {0.0, 0.5};
  s.iterations = 1000;
  s.algorithm = "pdhg-gpu";
  s.message = "stopped";
  s.primal_infeasibility_scaled = 1e-5;
  s.dual_infeasibility_scaled = 1e-6;
  return s;
}

}  // namespace

int main() {
  const Model model = two_by_one();
  Logger quiet(nullptr);
  const Timer timer;
  Options options;
  options.set_int("polish_iteration_limit", 7);

  // No engine installed: nothing happens.
  Solution s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 0 && s.algorithm == "pdhg-gpu" && s.iterations == 1000, "no engine, no polish");

  install_polish_engine(&stub);
  check(installed_polish_engine() == &stub, "engine installed");

  // An optimal answer, a status without a point, or vectors of the wrong size are not polished.
  s = pdhg_answer(SolveStatus::kOptimal);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 0, "optimal is not polished");
  s = pdhg_answer(SolveStatus::kNumericalError);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 0, "a numerical error holds no point to polish");
  s = pdhg_answer(SolveStatus::kIterationLimit);
  s.row_dual.clear();
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 0, "wrong-size vectors are not handed over");

  // pdhg_polish off: not polished.
  Options off = options;
  off.set_bool("pdhg_polish", false);
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, off, quiet, nullptr, timer);
  check(calls == 0, "pdhg_polish=false");

  // A polish that reaches optimal replaces the answer: the warm start is PDHG's point, the
  // engine gets polish_iteration_limit and polish_max_seconds (no time limit set), the
  // iterations add up and the algorithm reads pdhg-gpu+ipm.
  reply = Solution{};
  reply.status = SolveStatus::kOptimal;
  reply.iterations = 5;
  reply.objective = 1.75;
  reply.algorithm = "ipm";
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 1, "the engine ran once");
  check(seen.col_value == std::vector<double>({0.25, 0.75}) && seen.row_dual == std::vector<double>({-1.5}) &&
            seen.col_dual == std::vector<double>({0.0, 0.5}),
        "the warm start is PDHG's point, row duals and reduced costs");
  check(seen_iterations == 7, "iteration budget is polish_iteration_limit");
  check(seen_time == 30.0, "time budget is polish_max_seconds without a time limit");
  check(s.status == SolveStatus::kOptimal && s.objective == 1.75, "the polished answer replaced PDHG's");
  check(s.iterations == 1005 && s.polish_iterations == 5, "iterations are the sum of both phases");
  check(s.algorithm == "pdhg-gpu+ipm", "algorithm pdhg-gpu+ipm");
  check(s.message.find("polished by the interior point in 5 iterations") != std::string::npos,
        "the message says the polish ran");

  // A feasible polish with larger scaled violations does not replace it; its iterations are
  // recorded and the message says why.
  reply.status = SolveStatus::kFeasible;
  reply.iterations = 3;
  reply.primal_infeasibility_scaled = 1e-3;
  reply.message = "slow";
  s = pdhg_answer(SolveStatus::kFeasible);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(calls == 2, "the engine ran again");
  check(s.status == SolveStatus::kFeasible && s.algorithm == "pdhg-gpu" && s.iterations == 1000,
        "a worse polish leaves PDHG's answer");
  check(s.polish_iterations == 3, "its iterations are recorded");
  check(s.message.find("did not improve it") != std::string::npos, "the message says why");

  // A feasible polish with smaller violations replaces it.
  reply.primal_infeasibility_scaled = 1e-9;
  reply.dual_infeasibility_scaled = 1e-9;
  s = pdhg_answer(SolveStatus::kTimeLimit);
  detail::polish_with_the_interior_point(&s, model, options, quiet, nullptr, timer);
  check(s.status == SolveStatus::kFeasible && s.algorithm == "pdhg-gpu+ipm", "a better feasible polish replaces it");

  // With a finite time limit the budget is the smaller of polish_max_seconds and what is left;
  // with nothing left there is no polish; deterministic mode ignores the clock.
  Options timed = options;
  timed.set_double("time_limit", 1000.0);
  timed.set_double("polish_max_seconds", 12.5);
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, timed, quiet, nullptr, timer);
  check(seen_time == 12.5, "polish_max_seconds when the limit leaves more");
  timed.set_double("polish_max_seconds", 5000.0);
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, timed, quiet, nullptr, timer);
  check(seen_time > 990.0 && seen_time <= 1000.0, "what the time limit has left");
  const int before = calls;
  timed.set_double("time_limit", 0.0);
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, timed, quiet, nullptr, timer);
  check(calls == before && s.message.find("no time left") != std::string::npos, "no time left");
  timed.set_bool("deterministic", true);
  timed.set_double("polish_max_seconds", 12.5);
  s = pdhg_answer(SolveStatus::kIterationLimit);
  detail::polish_with_the_interior_point(&s, model, timed, quiet, nullptr, timer);
  check(calls == before + 1 && seen_time == 12.5, "deterministic: polish_max_seconds alone");

  install_polish_engine(nullptr);
  check(installed_polish_engine() == nullptr, "engine removed");
  if (failures == 0) std::printf("polish handoff: all checks passed\n");
  return failures == 0 ? 0 : 1;
}
