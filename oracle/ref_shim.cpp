// ref_shim.cpp - a C-ABI caller of the reference's own scaling, power iteration, SpMV,
// evaluate(), Solution::recompute_quality and Model::validate functions, compiled together with those reference sources where they lie
// (oracle/Makefile, target `ref`). It exists only to check the C restatement in
// pdhg_oracle.c against the reference bit for bit; nothing here re-implements the
// reference. Test infrastructure only.
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "la/scaling.hpp"
#include "pdhg/pdhg_evaluate.hpp"
#include "sankhya/model.hpp"

using namespace sankhya;

namespace {

Model make_model(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                 const double* val, const double* cost, const double* lo, const double* hi,
                 const double* rlo, const double* rhi) {
  Model model;
  model.col_cost.assign(cost, cost + n);
  model.col_lower.assign(lo, lo + n);
  model.col_upper.assign(hi, hi + n);
  model.col_type.assign(static_cast<std::size_t>(n), VarType::kContinuous);
  model.row_lower.assign(rlo, rlo + m);
  model.row_upper.assign(rhi, rhi + m);
  model.matrix.reset(m, n);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) model.matrix.add_entry(rowidx[k], j, val[k]);
  model.matrix.finalize(0.0);
  return model;
}

}  // namespace

extern "C" {

// Scaled CSC values come back in the reference's own order (rows ascending per column).
int ref_build_scaling(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                      const double* val, const double* cost, const double* lo, const double* hi,
                      const double* rlo, const double* rhi, const double* min_cost, int ruiz,
                      double* row, double* column, double* scaled_val, double* chat,
                      double* lhat, double* uhat, double* rlhat, double* ruhat) {
  const Model model = make_model(m, n, colptr, rowidx, val, cost, lo, hi, rlo, rhi);
  const std::vector<double> mc(min_cost, min_cost + n);
  const Scaling s = build_scaling(model, mc, ruiz);
  std::memcpy(row, s.row.data(), sizeof(double) * m);
  std::memcpy(column, s.column.data(), sizeof(double) * n);
  const auto& v = s.matrix.values();
  std::memcpy(scaled_val, v.data(), sizeof(double) * v.size());
  std::memcpy(chat, s.cost.data(), sizeof(double) * n);
  std::memcpy(lhat, s.col_lower.data(), sizeof(double) * n);
  std::memcpy(uhat, s.col_upper.data(), sizeof(double) * n);
  std::memcpy(rlhat, s.row_lower.data(), sizeof(double) * m);
  std::memcpy(ruhat, s.row_upper.data(), sizeof(double) * m);
  return static_cast<int>(v.size());
}

double ref_spectral_norm(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                         const double* val, int iterations, uint32_t seed) {
  SparseMatrix a(m, n);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) a.add_entry(rowidx[k], j, val[k]);
  a.finalize(0.0);
  return estimate_spectral_norm(a, iterations, seed);
}

// out[9] = primal, dual, gap, primal_objective, dual_objective, absolute_primal,
// absolute_dual, gap_as_verified, complementarity
void ref_evaluate(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                  const double* val, const double* cost, const double* lo, const double* hi,
                  const double* rlo, const double* rhi, const double* min_cost,
                  double bound_norm, double cost_norm, const double* x, const double* y,
                  double* activity, double* reduced, double* out) {
  const Model model = make_model(m, n, colptr, rowidx, val, cost, lo, hi, rlo, rhi);
  pdhg::Problem problem;
  problem.model = &model;
  problem.cost.assign(min_cost, min_cost + n);
  problem.bound_norm = bound_norm;
  problem.cost_norm = cost_norm;
  std::vector<double> xs(x, x + n), ys(y, y + m), act, red;
  const pdhg::Residuals r = pdhg::evaluate(problem, xs, ys, act, red);
  std::memcpy(activity, act.data(), sizeof(double) * m);
  std::memcpy(reduced, red.data(), sizeof(double) * n);
  const double vals[9] = {r.primal,          r.dual,          r.gap,
                          r.primal_objective, r.dual_objective, r.absolute_primal,
                          r.absolute_dual,   r.gap_as_verified, r.complementarity};
  std::memcpy(out, vals, sizeof(vals));
}

// Solution::recompute_quality (src/core/model.cpp:247-441) of the point (x, row_dual, col_dual)
// with the given dual bound on the model in its own sense. activity (m) out; out[9] =
// objective, primal_infeasibility, primal_infeasibility_scaled, dual_infeasibility,
// dual_infeasibility_scaled, complementarity_violation, integrality_violation, absolute_gap,
// relative_gap.
void ref_recompute_quality(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                           const double* val, const double* cost, const double* lo,
                           const double* hi, const double* rlo, const double* rhi, double offset,
                           int maximize, const double* x, const double* row_dual,
                           const double* col_dual, double dual_bound, double* activity,
                           double* out) {
  Model model = make_model(m, n, colptr, rowidx, val, cost, lo, hi, rlo, rhi);
  model.objective_offset = offset;
  model.sense = maximize ? ObjSense::kMaximize : ObjSense::kMinimize;
  Solution s;
  s.col_value.assign(x, x + n);
  s.row_dual.assign(row_dual, row_dual + m);
  s.col_dual.assign(col_dual, col_dual + n);
  s.dual_bound = dual_bound;
  s.recompute_quality(model);
  if (m > 0) std::memcpy(activity, s.row_activity.data(), sizeof(double) * m);
  const double vals[9] = {s.objective,          s.primal_infeasibility,
                          s.primal_infeasibility_scaled, s.dual_infeasibility,
                          s.dual_infeasibility_scaled,   s.complementarity_violation,
                          s.integrality_violation,       s.absolute_gap,
                          s.relative_gap};
  std::memcpy(out, vals, sizeof(vals));
}

// Model::validate (src/core/model.cpp:127-237) of the model as given (entries added one by one
// and finalised without a drop tolerance). The message goes to buf (empty when well formed).
void ref_validate(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                  const double* val, const double* cost, const double* lo, const double* hi,
                  const double* rlo, const double* rhi, char* buf, int len) {
  const Model model = make_model(m, n, colptr, rowidx, val, cost, lo, hi, rlo, rhi);
  const std::string text = model.validate();
  const std::size_t k = std::min<std::size_t>(static_cast<std::size_t>(len - 1), text.size());
  std::memcpy(buf, text.data(), k);
  buf[k] = '\0';
}

}  // extern "C"
