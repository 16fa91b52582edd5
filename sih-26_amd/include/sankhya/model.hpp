// SPDX-License-Identifier: Apache-2.0
// Model and Solution as the PDHG path reads and writes them (reference:
// include/sankhya/model.hpp:40-615). The frozen-interface fields that belong to other
// engines (basis ranging, IIS, presolve reports, solution pools) are not part of this path.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "sankhya/solve_control.hpp"
#include "sankhya/sparse.hpp"
#include "sankhya/tolerances.hpp"
#include "sankhya/types.hpp"

namespace sankhya {

class Options;

enum class ObjSense { kMinimize, kMaximize };
enum class VarType : std::uint8_t { kContinuous, kInteger };
enum class BasisStatus : std::uint8_t { kUnknown, kBasic, kAtLower, kAtUpper, kNonbasicFree, kFixed };

/// Terminal state of a solve; the order is the reference's (model.hpp:64-96).
enum class SolveStatus : std::uint8_t {
  kNotSolved,
  kOptimal,
  kInfeasible,
  kUnbounded,
  kInfeasibleOrUnbounded,
  kFeasible,
  kIterationLimit,
  kTimeLimit,
  kNodeLimit,
  kNumericalError,
  kModelError,
  kInterrupted
};

enum class LimitReason : std::uint8_t { kNone, kInterrupt, kTime, kIterations, kNodes };

[[nodiscard]] const char* to_string(SolveStatus status) noexcept;
[[nodiscard]] const char* to_string(LimitReason reason) noexcept;
[[nodiscard]] const char* to_string(BasisStatus status) noexcept;
[[nodiscard]] SolveStatus status_for(LimitReason reason) noexcept;

/// Does a solve ending in this state hand back a point? (reference model.hpp:581-597)
[[nodiscard]] constexpr bool claims_a_point(SolveStatus s) noexcept {
  switch (s) {
    case SolveStatus::kOptimal:
    case SolveStatus::kFeasible:
    case SolveStatus::kUnbounded:
    case SolveStatus::kIterationLimit:
    case SolveStatus::kTimeLimit:
    case SolveStatus::kNodeLimit:
    case SolveStatus::kInterrupted: return true;
    default: return false;
  }
}

///     optimize   objective_offset + c.x
///     subject to row_lower <= A x <= row_upper,  col_lower <= x <= col_upper
class Model {
 public:
  std::string name;
  std::string source_path;
  ObjSense sense = ObjSense::kMinimize;
  double objective_offset = 0.0;
  std::vector<double> col_cost;
  SparseMatrix hessian;
  std::vector<double> col_lower, col_upper;
  std::vector<VarType> col_type;
  std::vector<std::string> col_names;
  std::vector<double> row_lower, row_upper;
  std::vector<std::string> row_names;
  SparseMatrix matrix;

  [[nodiscard]] Index num_cols() const noexcept { return static_cast<Index>(col_cost.size()); }
  [[nodiscard]] Index num_rows() const noexcept { return static_cast<Index>(row_lower.size()); }
  [[nodiscard]] Index num_nonzeros() const noexcept { return matrix.num_nonzeros(); }
  [[nodiscard]] bool has_integrality() const noexcept;
  [[nodiscard]] bool has_quadratic_objective() const noexcept { return hessian.num_nonzeros() > 0; }
  [[nodiscard]] double sense_multiplier() const noexcept {
    return sense == ObjSense::kMaximize ? -1.0 : 1.0;
  }
  void resize_columns(Index n);
  void resize_rows(Index m);
  [[nodiscard]] double evaluate_objective(const double* x) const;
  /// Empty when well formed, else the first problem found (reference model.cpp:127-237).
  [[nodiscard]] std::string validate() const;
};

class Solution {
 public:
  SolveStatus status = SolveStatus::kNotSolved;
  LimitReason stopped_by = LimitReason::kNone;
  double objective = 0.0;
  double dual_bound = 0.0;
  std::vector<double> col_value, row_activity, row_dual, col_dual;
  std::vector<BasisStatus> col_status, row_status;

  double primal_infeasibility = 0.0;
  double primal_infeasibility_scaled = 0.0;
  double dual_infeasibility_scaled = 0.0;
  double dual_infeasibility = 0.0;
  double complementarity_violation = 0.0;
  double integrality_violation = 0.0;
  double absolute_gap = 0.0;
  double relative_gap = 0.0;

  Count iterations = 0;
  Count polish_iterations = 0;
  double solve_seconds = 0.0;
  std::string algorithm;
  std::string message;

  [[nodiscard]] bool has_primal_values() const noexcept { return claims_a_point(status); }
  void allocate_for(const Model& model);
  /// Row activities, infeasibilities and gaps measured from col_value / row_dual / col_dual
  /// (reference model.cpp:247-441).
  void recompute_quality(const Model& model);
};

/// The single entry point (reference model.hpp:609-613). In this build the LP branch is the
/// only one: it validates, selects the MI355X PDHG engine and runs it.
[[nodiscard]] Solution solve(const Model& model, const Options& options,
                             SolveControl* control = nullptr);

}  // namespace sankhya
