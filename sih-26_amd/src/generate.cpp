// SPDX-License-Identifier: Apache-2.0
// Deterministic synthetic LP generators for the benchmark and test workloads.
//
// Every instance is a function of (kind, sizes, seed) alone: one SplitMix64 stream feeds
// uniform draws (53-bit mantissas) and Box-Muller normals, so the same seed gives the same
// model on any host with the same libm. Kinds:
//   0  uniform-equality: `per_col` distinct uniformly random rows per column, values N(0,1),
//      every row an equality. x0 ~ U[0,1], bounds [0,10], b = A x0; y0 ~ N(0,1);
//      z ~ U[0,1] on a random half of the columns (exactly floor(n/2)), 0 elsewhere;
//      c = A^T y0 + z. Feasible (x0) and bounded (the box).
//   1  zipf: as kind 0, but the rows of each column are drawn from a bounded Zipf(param)
//      law over row ids (rejecting repeats), so row lengths follow a power law, and a random
//      30% of the rows are `<=` (row_lower = -inf, row_upper = b) instead of equalities.
//   2  blend: kind 0 on m-1 rows plus one dense row sum_j x_j = sum_j x0_j, which bounds the
//      feasible set; a random 40% of the columns get an upper bound U[1,50], the rest none;
//      c ~ U[-5,5].
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "generate.hpp"

namespace sankhya::hip_engine {
namespace {

struct Rng {
  std::uint64_t s;
  bool has_spare = false;
  double spare = 0.0;
  // The seed goes through the output mix before it becomes the state: the state advances by
  // the increment below, so seeding with an affine function of the seed would make seed k + 1
  // replay seed k's stream one draw later.
  explicit Rng(std::uint64_t seed) : s(mix(seed ^ 0x632BE59BD9B4E019ULL)) {}
  static std::uint64_t mix(std::uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  std::uint64_t next() { return mix(s += 0x9E3779B97F4A7C15ULL); }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double normal() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double t = 6.283185307179586 * u2;
    spare = r * std::sin(t);
    has_spare = true;
    return r * std::cos(t);
  }
  std::uint32_t below(std::uint32_t k) { return static_cast<std::uint32_t>(uniform() * k); }
};

// Vose alias table for the bounded Zipf law P(k) ~ (k+1)^-s, k = 0..m-1.
struct Alias {
  std::vector<double> prob;
  std::vector<std::uint32_t> alias;
  Alias(std::int32_t m, double s) : prob(static_cast<std::size_t>(m)), alias(static_cast<std::size_t>(m)) {
    std::vector<double> w(static_cast<std::size_t>(m));
    double total = 0.0;
    for (std::int32_t k = 0; k < m; ++k) total += (w[k] = std::pow(static_cast<double>(k + 1), -s));
    std::vector<std::uint32_t> small, large;
    for (std::int32_t k = 0; k < m; ++k) {
      w[k] = w[k] * m / total;
      (w[k] < 1.0 ? small : large).push_back(static_cast<std::uint32_t>(k));
    }
    while (!small.empty() && !large.empty()) {
      const std::uint32_t a = small.back(), b = large.back();
      small.pop_back();
      prob[a] = w[a];
      alias[a] = b;
      w[b] = (w[b] + w[a]) - 1.0;
      if (w[b] < 1.0) {
        large.pop_back();
        small.push_back(b);
      }
    }
    for (const std::uint32_t k : large) prob[k] = 1.0, alias[k] = k;
    for (const std::uint32_t k : small) prob[k] = 1.0, alias[k] = k;
  }
  std::uint32_t draw(Rng& r) const {
    const std::uint32_t k = r.below(static_cast<std::uint32_t>(prob.size()));
    return r.uniform() < prob[k] ? k : alias[k];
  }
};

// Exactly `count` distinct indices out of [0, total), by a partial Fisher-Yates shuffle.
std::vector<char> pick(Rng& r, std::int32_t total, std::int32_t count) {
  std::vector<std::int32_t> perm(static_cast<std::size_t>(total));
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<char> chosen(static_cast<std::size_t>(total), 0);
  for (std::int32_t k = 0; k < count && k < total; ++k) {
    const std::int32_t t = k + static_cast<std::int32_t>(r.below(static_cast<std::uint32_t>(total - k)));
    std::swap(perm[k], perm[t]);
    chosen[perm[k]] = 1;
  }
  return chosen;
}

}  // namespace

Model generate(const GenSpec& spec) {
  const std::int32_t m = spec.m, n = spec.n;
  const bool blend = spec.kind == 2;
  const std::int32_t mr = blend ? m - 1 : m;  // rows filled at random
  const std::int32_t per = std::min(spec.per_col, std::max(mr, 0));
  Rng rng(spec.seed ^ (static_cast<std::uint64_t>(spec.kind) << 56));
  Model model;
  model.name = spec.kind == 1 ? "zipf" : (blend ? "blend" : "uniform");
  model.resize_columns(n);
  model.resize_rows(m);

  const Alias* zipf = nullptr;
  Alias* zipf_table = spec.kind == 1 ? new Alias(mr, spec.param) : nullptr;
  zipf = zipf_table;

  std::vector<Index> starts(static_cast<std::size_t>(n) + 1, 0);
  std::vector<Index> rows;
  std::vector<double> vals;
  const std::size_t nnz_guess = static_cast<std::size_t>(n) * static_cast<std::size_t>(per + (blend ? 1 : 0));
  rows.reserve(nnz_guess);
  vals.reserve(nnz_guess);
  std::vector<std::pair<Index, double>> col;
  for (std::int32_t j = 0; j < n; ++j) {
    col.clear();
    while (static_cast<std::int32_t>(col.size()) < per) {
      const Index i = zipf ? static_cast<Index>(zipf->draw(rng)) : static_cast<Index>(rng.below(static_cast<std::uint32_t>(mr)));
      bool dup = false;
      for (const auto& e : col) dup = dup || e.first == i;
      if (!dup) col.push_back({i, 0.0});
    }
    for (auto& e : col) e.second = rng.normal();
    if (blend) col.push_back({m - 1, 1.0});
    std::sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& e : col) {
      rows.push_back(e.first);
      vals.push_back(e.second);
    }
    starts[j + 1] = static_cast<Index>(rows.size());
  }
  delete zipf_table;

  std::vector<double> x0(static_cast<std::size_t>(n));
  for (double& v : x0) v = rng.uniform();
  std::vector<double> b(static_cast<std::size_t>(m), 0.0);
  for (std::int32_t j = 0; j < n; ++j)
    for (Index k = starts[j]; k < starts[j + 1]; ++k) b[rows[k]] += vals[k] * x0[j];

  for (std::int32_t j = 0; j < n; ++j) {
    model.col_lower[j] = 0.0;
    model.col_upper[j] = blend ? kInfinity : 10.0;
  }
  if (blend) {
    const std::vector<char> bounded = pick(rng, n, static_cast<std::int32_t>(std::llround(0.4 * n)));
    for (std::int32_t j = 0; j < n; ++j)
      if (bounded[j]) model.col_upper[j] = 1.0 + 49.0 * rng.uniform();
  }
  std::vector<char> leq(static_cast<std::size_t>(m), 0);
  if (spec.kind == 1) leq = pick(rng, m, static_cast<std::int32_t>(std::llround(0.3 * m)));
  for (std::int32_t i = 0; i < m; ++i) {
    model.row_upper[i] = b[i];
    model.row_lower[i] = leq[i] ? -kInfinity : b[i];
  }
  if (blend) {
    for (std::int32_t j = 0; j < n; ++j) model.col_cost[j] = -5.0 + 10.0 * rng.uniform();
  } else {
    std::vector<double> y0(static_cast<std::size_t>(m));
    for (double& v : y0) v = rng.normal();
    const std::vector<char> half = pick(rng, n, n / 2);
    for (std::int32_t j = 0; j < n; ++j) {
      double dot = 0.0;
      for (Index k = starts[j]; k < starts[j + 1]; ++k) dot += vals[k] * y0[rows[k]];
      model.col_cost[j] = dot + (half[j] ? rng.uniform() : 0.0);
    }
  }
  model.matrix.assign_csc(m, n, std::move(starts), std::move(rows), std::move(vals));
  return model;
}

}  // namespace sankhya::hip_engine
