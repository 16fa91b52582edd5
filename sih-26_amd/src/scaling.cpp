// SPDX-License-Identifier: Apache-2.0
// Host preconditioner (see scaling.hpp). Reference: src/la/scaling.cpp:14-159.
#include "scaling.hpp"

#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <type_traits>
#include <vector>

namespace sankhya::hip_engine {
namespace {

double scale_bound(double bound, double multiplier) {
  return is_finite_bound(bound) ? bound * multiplier : bound;
}

}  // namespace


Scaling build_scaling(const Model& model, const std::vector<double>& min_space_cost,
                      int ruiz_iterations, int threads) {
  const Index m = model.num_rows(), n = model.num_cols();
  const auto& starts = model.matrix.column_starts();
  const auto& rows = model.matrix.row_indices();
  const std::size_t um = static_cast<std::size_t>(m), un = static_cast<std::size_t>(n);
  // Small models are not worth the threads.
  if (model.num_nonzeros() < 200000) threads = 1;
  threads = std::max(1, threads);
  const bool split = threads > 1;
  Scaling s;
  s.row.assign(um, 1.0);
  s.column.assign(un, 1.0);
  s.values = model.matrix.values();
  std::vector<double> rstat(um), cstat(un), rpass(um), cpass(un);
  // Per-thread row maxima (Ruiz): a maximum does not depend on the order it is taken in, so
  // merging the threads' maxima gives the sequential answer bit for bit.
  std::vector<std::vector<double>> rmax(static_cast<std::size_t>(split ? threads : 0),
                                        std::vector<double>(um));
  const std::size_t chunk_cols = std::max<std::size_t>(1, split ? (un + threads - 1) / threads : un);

  // One sweep over the columns: first scale each entry by this pass's factors,
  // (a * r_i) * c_j (scaling.cpp:14-40), then, when the next pass is a Ruiz pass, take the
  // infinity norms it needs from the scaled entries (scaling.cpp:50-59). Fusing the two gives
  // the same values and maxima as separate passes.
  auto sweep = [&](bool apply, bool maxima) {
    const Index* st = starts.data();
    const Index* rw = rows.data();
    const double* rp = rpass.data();
    const double* cp = cpass.data();
    double* vals = s.values.data();
    double* cs = cstat.data();
    auto body = [&](auto apply_c, auto maxima_c, std::size_t b, std::size_t e, double* rt) {
      constexpr bool kApply = decltype(apply_c)::value, kMaxima = decltype(maxima_c)::value;
      if (kMaxima) std::fill(rt, rt + um, 0.0);
      for (std::size_t j = b; j < e; ++j) {
        const double cj = cp[j];
        double c = 0.0;
        for (Index k = st[j], k1 = st[j + 1]; k < k1; ++k) {
          double v = vals[k];
          if (kApply) {
            v = v * rp[rw[k]] * cj;
            vals[k] = v;
          }
          if (kMaxima) {
            const double mag = std::fabs(v);
            double& r = rt[rw[k]];
            r = std::max(r, mag);
            c = std::max(c, mag);
          }
        }
        if (kMaxima) cs[j] = c;
      }
    };
    auto run = [&](std::size_t b, std::size_t e, double* rt) {
      using T = std::true_type;
      using F = std::false_type;
      if (apply && maxima) body(T{}, T{}, b, e, rt);
      else if (apply) body(T{}, F{}, b, e, rt);
      else body(F{}, T{}, b, e, rt);
    };
    if (!split) run(0, un, rstat.data());
    else parallel_ranges(un, threads, [&](std::size_t b, std::size_t e) { run(b, e, rmax[b / chunk_cols].data()); });
  };
  // This pass's factors r_i = 1/sqrt(stat_i), c_j = 1/sqrt(stat_j) (1 for an empty line) and
  // the running products Dr *= r, Dc *= c; `merged` takes the row statistic from the
  // per-thread maxima.
  auto factors = [&](bool merged) {
    parallel_ranges(std::max(um, un), threads, [&](std::size_t b, std::size_t e) {
      for (std::size_t i = b; i < std::min(e, um); ++i) {
        double r = rstat[i];
        if (merged) {
          r = 0.0;
          for (const std::vector<double>& rt : rmax) r = std::max(r, rt[i]);
        }
        rpass[i] = r > 0.0 ? 1.0 / std::sqrt(r) : 1.0;
        s.row[i] *= rpass[i];
      }
      for (std::size_t j = b; j < std::min(e, un); ++j) {
        cpass[j] = cstat[j] > 0.0 ? 1.0 / std::sqrt(cstat[j]) : 1.0;
        s.column[j] *= cpass[j];
      }
    });
  };

  if (ruiz_iterations > 0) sweep(false, true);
  for (int pass = 0; pass < ruiz_iterations; ++pass) {
    factors(split);
    sweep(true, pass + 1 < ruiz_iterations);
  }
  // Pock-Chambolle (alpha = 1): 1-norm sums (scaling.cpp:76-100). A sum depends on the order of
  // its terms, so this pass stays one sequential sweep in CSC order, as in the reference.
  std::fill(rstat.begin(), rstat.end(), 0.0);
  for (Index j = 0; j < n; ++j) {
    double c = 0.0;
    for (Index k = starts[j]; k < starts[j + 1]; ++k) {
      const double mag = std::fabs(s.values[k]);
      rstat[rows[k]] += mag;
      c += mag;
    }
    cstat[j] = c;
  }
  factors(false);
  sweep(true, false);

  s.cost.resize(un);
  s.col_lower.resize(un);
  s.col_upper.resize(un);
  s.row_lower.resize(um);
  s.row_upper.resize(um);
  parallel_ranges(std::max(um, un), threads, [&](std::size_t b, std::size_t e) {
    for (std::size_t j = b; j < std::min(e, un); ++j) {
      const double dc = s.column[j];
      s.cost[j] = min_space_cost[j] * dc;
      s.col_lower[j] = scale_bound(model.col_lower[j], 1.0 / dc);
      s.col_upper[j] = scale_bound(model.col_upper[j], 1.0 / dc);
    }
    for (std::size_t i = b; i < std::min(e, um); ++i) {
      const double dr = s.row[i];
      s.row_lower[i] = scale_bound(model.row_lower[i], dr);
      s.row_upper[i] = scale_bound(model.row_upper[i], dr);
    }
  });
  // Extremes of |Ahat| (minimum and maximum are exact in any order).
  const std::size_t nnz = s.values.size();
  if (nnz > 0) {
    const std::size_t t = static_cast<std::size_t>(threads);
    std::vector<double> lo(t, std::fabs(s.values[0])), hi(t, 0.0);
    const std::size_t chunk = (nnz + t - 1) / t;
    parallel_ranges(nnz, threads, [&](std::size_t b, std::size_t e) {
      const std::size_t w = b / chunk;
      double a = lo[w], z = hi[w];
      for (std::size_t k = b; k < e; ++k) {
        const double mag = std::fabs(s.values[k]);
        a = std::min(a, mag);
        z = std::max(z, mag);
      }
      lo[w] = a;
      hi[w] = z;
    });
    s.min_abs = *std::min_element(lo.begin(), lo.end());
    s.max_abs = *std::max_element(hi.begin(), hi.end());
  }
  return s;
}

}  // namespace sankhya::hip_engine
