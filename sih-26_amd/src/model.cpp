// SPDX-License-Identifier: Apache-2.0
// Model / Solution behaviour the PDHG path relies on (reference: src/core/model.cpp:100-441).
#include <algorithm>
#include <cmath>
#include <string>

#include "sankhya/logging.hpp"
#include "sankhya/model.hpp"

namespace sankhya {

const char* to_string(SolveStatus s) noexcept {
  switch (s) {
    case SolveStatus::kNotSolved: return "not_solved";
    case SolveStatus::kOptimal: return "optimal";
    case SolveStatus::kInfeasible: return "infeasible";
    case SolveStatus::kUnbounded: return "unbounded";
    case SolveStatus::kInfeasibleOrUnbounded: return "infeasible_or_unbounded";
    case SolveStatus::kFeasible: return "feasible";
    case SolveStatus::kIterationLimit: return "iteration_limit";
    case SolveStatus::kTimeLimit: return "time_limit";
    case SolveStatus::kNodeLimit: return "node_limit";
    case SolveStatus::kNumericalError: return "numerical_error";
    case SolveStatus::kModelError: return "model_error";
    case SolveStatus::kInterrupted: return "interrupted";
  }
  return "unknown";
}

const char* to_string(LimitReason r) noexcept {
  switch (r) {
    case LimitReason::kNone: return "none";
    case LimitReason::kInterrupt: return "user_interrupt";
    case LimitReason::kTime: return "time_limit";
    case LimitReason::kIterations: return "iteration_limit";
    case LimitReason::kNodes: return "node_limit";
  }
  return "none";
}

const char* to_string(BasisStatus b) noexcept {
  switch (b) {
    case BasisStatus::kUnknown: return "unknown";
    case BasisStatus::kBasic: return "basic";
    case BasisStatus::kAtLower: return "at_lower";
    case BasisStatus::kAtUpper: return "at_upper";
    case BasisStatus::kNonbasicFree: return "free";
    case BasisStatus::kFixed: return "fixed";
  }
  return "unknown";
}

SolveStatus status_for(LimitReason r) noexcept {
  switch (r) {
    case LimitReason::kInterrupt: return SolveStatus::kInterrupted;
    case LimitReason::kTime: return SolveStatus::kTimeLimit;
    case LimitReason::kIterations: return SolveStatus::kIterationLimit;
    case LimitReason::kNodes: return SolveStatus::kNodeLimit;
    case LimitReason::kNone: break;
  }
  return SolveStatus::kNotSolved;
}

bool Model::has_integrality() const noexcept {
  return std::any_of(col_type.begin(), col_type.end(), [](VarType t) { return t == VarType::kInteger; });
}

void Model::resize_columns(Index n) {
  const auto u = static_cast<std::size_t>(n);
  col_cost.resize(u, 0.0);
  col_lower.resize(u, 0.0);
  col_upper.resize(u, kInfinity);
  col_type.resize(u, VarType::kContinuous);
  if (!col_names.empty()) col_names.resize(u);
}

void Model::resize_rows(Index m) {
  const auto u = static_cast<std::size_t>(m);
  row_lower.resize(u, -kInfinity);
  row_upper.resize(u, kInfinity);
  if (!row_names.empty()) row_names.resize(u);
}

double Model::evaluate_objective(const double* x) const {
  double linear = objective_offset;
  for (Index j = 0; j < num_cols(); ++j) linear += col_cost[j] * x[j];
  return linear;
}

std::string Model::validate() const {
  const Index n = num_cols(), m = num_rows();
  if (static_cast<Index>(col_lower.size()) != n) return "col_lower length != num_cols";
  if (static_cast<Index>(col_upper.size()) != n) return "col_upper length != num_cols";
  if (static_cast<Index>(col_type.size()) != n) return "col_type length != num_cols";
  if (!col_names.empty() && static_cast<Index>(col_names.size()) != n)
    return "col_names is neither empty nor num_cols long";
  if (static_cast<Index>(row_upper.size()) != m) return "row_upper length != num_rows";
  if (!row_names.empty() && static_cast<Index>(row_names.size()) != m)
    return "row_names is neither empty nor num_rows long";
  if (matrix.overflowed()) return "the constraint matrix exceeds the nonzero limit this build supports";
  if (!matrix.frozen()) return "constraint matrix is not finalized";
  if (matrix.num_rows() != m)
    return detail::format("matrix has {} rows, model has {}", matrix.num_rows(), m);
  if (matrix.num_cols() != n)
    return detail::format("matrix has {} columns, model has {}", matrix.num_cols(), n);
  for (Index j = 0; j < n; ++j) {
    const ColumnView c = matrix.column(j);
    for (Index k = 0; k < c.size; ++k)
      if (!std::isfinite(c.values[k]))
        return detail::format("matrix entry ({}, {}) is {}, which is not a usable coefficient",
                              c.rows[k], j, c.values[k]);
  }
  for (Index j = 0; j < n; ++j) {
    const double lo = col_lower[j], hi = col_upper[j];
    if (std::isnan(lo) || std::isnan(hi)) return detail::format("column {} has a NaN bound", j);
    if (lo > hi) return detail::format("column {} has lower bound {:g} above upper bound {:g}", j, lo, hi);
    if (std::isnan(col_cost[j])) return detail::format("column {} has a NaN objective coefficient", j);
  }
  for (Index i = 0; i < m; ++i) {
    const double lo = row_lower[i], hi = row_upper[i];
    if (std::isnan(lo) || std::isnan(hi)) return detail::format("row {} has a NaN bound", i);
    if (lo > hi) return detail::format("row {} has lower bound {:g} above upper bound {:g}", i, lo, hi);
  }
  if (std::isnan(objective_offset)) return "objective offset is NaN";
  return {};
}

void Solution::allocate_for(const Model& model) {
  const auto n = static_cast<std::size_t>(model.num_cols());
  const auto m = static_cast<std::size_t>(model.num_rows());
  col_value.assign(n, 0.0);
  col_dual.assign(n, 0.0);
  col_status.assign(n, BasisStatus::kUnknown);
  row_activity.assign(m, 0.0);
  row_dual.assign(m, 0.0);
  row_status.assign(m, BasisStatus::kUnknown);
}

void Solution::recompute_quality(const Model& model) {
  const Index n = model.num_cols(), m = model.num_rows();
  if (static_cast<Index>(col_value.size()) != n) return;
  row_activity.assign(static_cast<std::size_t>(m), 0.0);
  if (m > 0) model.matrix.multiply(col_value.data(), row_activity.data());
  primal_infeasibility = 0.0;
  primal_infeasibility_scaled = 0.0;
  integrality_violation = 0.0;
  std::vector<double> row_scale(static_cast<std::size_t>(m), 1.0);
  for (Index j = 0; j < n; ++j) {
    const double x = col_value[j];
    if (x == 0.0) continue;
    const ColumnView c = model.matrix.column(j);
    for (Index k = 0; k < c.size; ++k)
      row_scale[c.rows[k]] = std::max(row_scale[c.rows[k]], std::fabs(c.values[k] * x));
  }
  const auto record = [&](double violation, double scale) {
    if (violation <= 0.0) return;
    primal_infeasibility = std::max(primal_infeasibility, violation);
    primal_infeasibility_scaled = std::max(primal_infeasibility_scaled, violation / std::max(1.0, scale));
  };
  for (Index j = 0; j < n; ++j) {
    const double x = col_value[j];
    if (is_finite_bound(model.col_lower[j])) record(model.col_lower[j] - x, std::fabs(x));
    if (is_finite_bound(model.col_upper[j])) record(x - model.col_upper[j], std::fabs(x));
    if (model.col_type[j] == VarType::kInteger)
      integrality_violation = std::max(integrality_violation, std::fabs(x - std::round(x)));
  }
  for (Index i = 0; i < m; ++i) {
    const double a = row_activity[i];
    if (is_finite_bound(model.row_lower[i])) record(model.row_lower[i] - a, row_scale[i]);
    if (is_finite_bound(model.row_upper[i])) record(a - model.row_upper[i], row_scale[i]);
  }
  objective = model.evaluate_objective(col_value.data());

  if (static_cast<Index>(row_dual.size()) == m && static_cast<Index>(col_dual.size()) == n) {
    dual_infeasibility = 0.0;
    dual_infeasibility_scaled = 0.0;
    complementarity_violation = 0.0;
    const double sense = model.sense_multiplier();
    double dual_norm = 0.0;
    for (Index i = 0; i < m; ++i) dual_norm = std::max(dual_norm, std::fabs(row_dual[i]));
    const auto record_dual = [&](double violation, double scale) {
      if (violation <= 0.0) return;
      dual_infeasibility = std::max(dual_infeasibility, violation);
      dual_infeasibility_scaled = std::max(dual_infeasibility_scaled, violation / scale);
    };
    const auto nearest = [](double value, double lo, double hi) {
      double d = kInfinity;
      if (is_finite_bound(lo)) d = std::min(d, std::fabs(value - lo));
      if (is_finite_bound(hi)) d = std::min(d, std::fabs(value - hi));
      return d;
    };
    const auto record_compl = [&](double mult, double slack, double mscale, double pscale) {
      if (mult == 0.0 || !is_finite_bound(slack)) return;
      const double product = mult * slack;
      complementarity_violation = std::max(complementarity_violation, product);
      dual_infeasibility_scaled =
          std::max(dual_infeasibility_scaled, product / std::max(1.0, mscale * pscale));
    };
    const bool linear = !model.has_quadratic_objective() && !model.has_integrality();
    std::vector<double> col_scale(static_cast<std::size_t>(n), 1.0);
    for (Index j = 0; j < n; ++j) {
      double implied = model.col_cost[j];
      double terms = std::fabs(model.col_cost[j]);
      const ColumnView c = model.matrix.column(j);
      for (Index k = 0; k < c.size; ++k) {
        const double term = c.values[k] * row_dual[c.rows[k]];
        implied -= term;
        terms = std::max(terms, std::fabs(term));
      }
      col_scale[j] = std::max(1.0, terms);
      if (linear) record_dual(std::fabs(implied - col_dual[j]), std::max(1.0, terms));
    }
    for (Index j = 0; j < n; ++j) {
      const double d = sense * col_dual[j];
      const double x = col_value[j];
      const double lo = model.col_lower[j], hi = model.col_upper[j];
      if (lo == hi) continue;
      const double scale = col_scale[j];
      if (d > 0.0 && !is_finite_bound(lo)) record_dual(d, scale);
      if (d < 0.0 && !is_finite_bound(hi)) record_dual(-d, scale);
      record_compl(std::fabs(d), nearest(x, lo, hi), scale, std::max(1.0, std::fabs(x)));
    }
    for (Index i = 0; i < m; ++i) {
      const double y = sense * row_dual[i];
      const double a = row_activity[i];
      const double lo = model.row_lower[i], hi = model.row_upper[i];
      if (lo == hi) continue;
      const double price_scale = std::max(1.0, dual_norm);
      if (y > 0.0 && !is_finite_bound(lo)) record_dual(y, price_scale);
      if (y < 0.0 && !is_finite_bound(hi)) record_dual(-y, price_scale);
      record_compl(std::fabs(y), nearest(a, lo, hi), price_scale, row_scale[i]);
    }
  }
  absolute_gap = std::fabs(objective - dual_bound);
  relative_gap = absolute_gap / std::max(1.0, std::fabs(objective));
}

}  // namespace sankhya
