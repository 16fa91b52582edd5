// SPDX-License-Identifier: Apache-2.0
// Host preconditioner (see scaling.hpp). Reference: src/la/scaling.cpp:14-159.
#include "scaling.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace sankhya::hip_engine {
namespace {

double scale_bound(double bound, double multiplier) {
  return is_finite_bound(bound) ? bound * multiplier : bound;
}

}  // namespace

namespace {

// f(begin, end) over [0, count) cut into `threads` contiguous ranges, one per thread.
template <class F>
void parallel_ranges(std::size_t count, int threads, F&& f) {
  if (threads <= 1 || count < 2) {
    f(std::size_t{0}, count);
    return;
  }
  const std::size_t t = static_cast<std::size_t>(threads);
  const std::size_t chunk = (count + t - 1) / t;
  std::vector<std::thread> pool;
  for (std::size_t k = 1; k < t; ++k) {
    const std::size_t b = k * chunk, e = std::min(count, b + chunk);
    if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
  }
  f(std::size_t{0}, std::min(count, chunk));
  for (std::thread& th : pool) th.join();
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
  Scaling s;
  s.row.assign(um, 1.0);
  s.column.assign(un, 1.0);
  s.values = model.matrix.values();
  std::vector<double> rstat(um), cstat(un), rpass(um), cpass(un);
  // Per-thread row maxima (Ruiz): a maximum does not depend on the order it is taken in, so
  // merging the threads' maxima gives the sequential answer bit for bit.
  std::vector<std::vector<double>> rmax(static_cast<std::size_t>(threads > 1 ? threads : 0),
                                        std::vector<double>(um));
  // The Pock-Chambolle sums do depend on order, so that one pass stays a single sequential CSC
  // sweep, as in the reference.
  const std::size_t chunk_cols =
      std::max<std::size_t>(1, threads > 1 ? (un + threads - 1) / threads : un);
  for (int pass = 0; pass <= ruiz_iterations; ++pass) {
    const bool pock_chambolle = pass == ruiz_iterations;
    if (threads == 1 || pock_chambolle) {
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
    } else {
      parallel_ranges(un, threads, [&](std::size_t b, std::size_t e) {
        std::vector<double>& rt = rmax[b / chunk_cols];
        std::fill(rt.begin(), rt.end(), 0.0);
        for (std::size_t j = b; j < e; ++j) {
          double c = 0.0;
          for (Index k = starts[j]; k < starts[j + 1]; ++k) {
            const double mag = std::fabs(s.values[k]);
            rt[rows[k]] = std::max(rt[rows[k]], mag);
            c = std::max(c, mag);
          }
          cstat[j] = c;
        }
      });
      parallel_ranges(um, threads, [&](std::size_t b, std::size_t e) {
        for (std::size_t i = b; i < e; ++i) {
          double r = 0.0;
          for (const std::vector<double>& rt : rmax) r = std::max(r, rt[i]);
          rstat[i] = r;
        }
      });
    }
    for (Index i = 0; i < m; ++i) rpass[i] = rstat[i] > 0.0 ? 1.0 / std::sqrt(rstat[i]) : 1.0;
    for (Index j = 0; j < n; ++j) cpass[j] = cstat[j] > 0.0 ? 1.0 / std::sqrt(cstat[j]) : 1.0;
    parallel_ranges(un, threads, [&](std::size_t b, std::size_t e) {
      for (std::size_t j = b; j < e; ++j) {
        const double cj = cpass[j];
        for (Index k = starts[j]; k < starts[j + 1]; ++k)
          s.values[k] = s.values[k] * rpass[rows[k]] * cj;
      }
    });
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
