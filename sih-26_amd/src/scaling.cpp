// SPDX-License-Identifier: Apache-2.0
// Host preconditioner (see scaling.hpp). Reference: src/la/scaling.cpp:14-159.
#include "scaling.hpp"

#include <algorithm>
#include <cmath>

namespace sankhya::hip_engine {
namespace {

double scale_bound(double bound, double multiplier) {
  return is_finite_bound(bound) ? bound * multiplier : bound;
}

}  // namespace

Scaling build_scaling(const Model& model, const std::vector<double>& min_space_cost,
                      int ruiz_iterations) {
  const Index m = model.num_rows(), n = model.num_cols();
  const auto& starts = model.matrix.column_starts();
  const auto& rows = model.matrix.row_indices();
  Scaling s;
  s.row.assign(static_cast<std::size_t>(m), 1.0);
  s.column.assign(static_cast<std::size_t>(n), 1.0);
  s.values = model.matrix.values();
  std::vector<double> rstat(static_cast<std::size_t>(m)), cstat(static_cast<std::size_t>(n));
  std::vector<double> rpass(static_cast<std::size_t>(m)), cpass(static_cast<std::size_t>(n));
  for (int pass = 0; pass <= ruiz_iterations; ++pass) {
    const bool pock_chambolle = pass == ruiz_iterations;
    std::fill(rstat.begin(), rstat.end(), 0.0);
    for (Index j = 0; j < n; ++j) {
      double c = 0.0;
      for (Index k = starts[j]; k < starts[j + 1]; ++k) {
        const double mag = std::fabs(s.values[k]);
        double& r = rstat[rows[k]];
        if (pock_chambolle) {
          r += mag;  // Pock-Chambolle: 1-norm sums (scaling.cpp:82-89)
          c += mag;
        } else {
          r = std::max(r, mag);  // Ruiz: infinity norms (scaling.cpp:50-59)
          c = std::max(c, mag);
        }
      }
      cstat[j] = c;
    }
    for (Index i = 0; i < m; ++i) rpass[i] = rstat[i] > 0.0 ? 1.0 / std::sqrt(rstat[i]) : 1.0;
    for (Index j = 0; j < n; ++j) cpass[j] = cstat[j] > 0.0 ? 1.0 / std::sqrt(cstat[j]) : 1.0;
    for (Index j = 0; j < n; ++j) {
      const double cj = cpass[j];
      for (Index k = starts[j]; k < starts[j + 1]; ++k) s.values[k] = s.values[k] * rpass[rows[k]] * cj;
    }
    for (Index i = 0; i < m; ++i) s.row[i] *= rpass[i];
    for (Index j = 0; j < n; ++j) s.column[j] *= cpass[j];
  }
  s.cost.resize(static_cast<std::size_t>(n));
  s.col_lower.resize(static_cast<std::size_t>(n));
  s.col_upper.resize(static_cast<std::size_t>(n));
  for (Index j = 0; j < n; ++j) {
    const double dc = s.column[j];
    s.cost[j] = min_space_cost[j] * dc;
    s.col_lower[j] = scale_bound(model.col_lower[j], 1.0 / dc);
    s.col_upper[j] = scale_bound(model.col_upper[j], 1.0 / dc);
  }
  s.row_lower.resize(static_cast<std::size_t>(m));
  s.row_upper.resize(static_cast<std::size_t>(m));
  for (Index i = 0; i < m; ++i) {
    const double dr = s.row[i];
    s.row_lower[i] = scale_bound(model.row_lower[i], dr);
    s.row_upper[i] = scale_bound(model.row_upper[i], dr);
  }
  bool first = true;
  for (const double v : s.values) {
    const double mag = std::fabs(v);
    s.max_abs = std::max(s.max_abs, mag);
    if (first || mag < s.min_abs) {
      s.min_abs = mag;
      first = false;
    }
  }
  return s;
}

}  // namespace sankhya::hip_engine
