// kkt_instances.cpp - writes the 120 small LPs of the reference's simplex-agreement test
// (tests/unit/test_pdhg.cpp:178-222: std::mt19937_64 seeded 20260904, at most 6 rows and 6
// columns) as JSON, each with the optimum its construction plants.
//
// The instances are fixed by the order of the draws, so this program makes the same draws in
// the same order as the reference's generator (tests/oracles/lp_generator.cpp:10-48 for the
// dimension, coefficient and coin draws and the matrix fill, :122-184 for the KKT
// construction) with the same std:: distributions. libstdc++ defines how those distributions
// consume the engine; built with the same library, the stream matches the reference's.
//
// Construction: a point x* >= 0, multipliers y* >= 0 on a random half of the rows and reduced
// costs d >= 0 with d_j x*_j = 0; c = A'y* + d, rows A_i x >= b_i tight at x* where y*_i may
// be nonzero and slack elsewhere, upper bounds no tighter than x*. Then (x*, y*, d) satisfies
// the KKT conditions of min c'x, A x >= b, 0 <= x <= u, so c'x* is the optimum.
//
// Build and run (tests/golden/make_kkt_instances.sh):
//   g++ -O2 -std=c++17 kkt_instances.cpp -o kkt_instances && ./kkt_instances > kkt_20260904.json
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

namespace {

using Rng = std::mt19937_64;

struct Config {
  int min_rows = 2, max_rows = 6, min_cols = 2, max_cols = 6;
  double density = 0.6;
  int64_t magnitude = 6;
  double bounded = 0.35;
};

int draw_int(Rng& g, int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(g); }
int64_t draw_i64(Rng& g, int64_t lo, int64_t hi) {
  return std::uniform_int_distribution<int64_t>(lo, hi)(g);
}
bool coin(Rng& g, double p) { return std::uniform_real_distribution<double>(0.0, 1.0)(g) < p; }

struct Lp {
  int rows = 0, cols = 0;
  std::vector<std::vector<int64_t>> a;
  std::vector<int64_t> b, c, upper, x;  // upper < 0: no upper bound
  int64_t objective = 0;
};

Lp make(Rng& g, const Config& k) {
  Lp lp;
  lp.rows = draw_int(g, k.min_rows, k.max_rows);
  lp.cols = draw_int(g, k.min_cols, k.max_cols);
  const size_t m = static_cast<size_t>(lp.rows), n = static_cast<size_t>(lp.cols);
  lp.a.assign(m, std::vector<int64_t>(n, 0));
  lp.b.assign(m, 0);
  lp.c.assign(n, 0);
  lp.upper.assign(n, -1);
  // Matrix: each entry present with probability `density`; a row left all zero gets one
  // entry, forced nonzero.
  for (size_t i = 0; i < m; ++i) {
    bool nonzero = false;
    for (size_t j = 0; j < n; ++j) {
      if (!coin(g, k.density)) continue;
      lp.a[i][j] = draw_i64(g, -k.magnitude, k.magnitude);
      nonzero = nonzero || lp.a[i][j] != 0;
    }
    if (!nonzero) {
      const size_t j = static_cast<size_t>(draw_int(g, 0, lp.cols - 1));
      const int64_t v = draw_i64(g, -k.magnitude, k.magnitude);
      lp.a[i][j] = v != 0 ? v : 1;
    }
  }
  // Planted primal point, multipliers on the active rows, reduced costs.
  lp.x.assign(n, 0);
  for (size_t j = 0; j < n; ++j)
    if (coin(g, 0.6)) lp.x[j] = draw_i64(g, 0, 5);
  std::vector<int64_t> y(m, 0);
  std::vector<bool> active(m, false);
  for (size_t i = 0; i < m; ++i) {
    if (coin(g, 0.5)) {
      active[i] = true;
      y[i] = draw_i64(g, 0, 5);
    }
  }
  std::vector<int64_t> d(n, 0);
  for (size_t j = 0; j < n; ++j) d[j] = lp.x[j] > 0 ? 0 : draw_i64(g, 0, 5);
  for (size_t j = 0; j < n; ++j) {
    int64_t cj = d[j];
    for (size_t i = 0; i < m; ++i) cj += lp.a[i][j] * y[i];
    lp.c[j] = cj;
  }
  for (size_t i = 0; i < m; ++i) {
    int64_t act = 0;
    for (size_t j = 0; j < n; ++j) act += lp.a[i][j] * lp.x[j];
    lp.b[i] = active[i] ? act : act - 1 - draw_i64(g, 0, 5);
  }
  for (size_t j = 0; j < n; ++j)
    if (coin(g, k.bounded)) lp.upper[j] = lp.x[j] + draw_i64(g, 0, 3);
  for (size_t j = 0; j < n; ++j) lp.objective += lp.c[j] * lp.x[j];
  return lp;
}

void list(const std::vector<int64_t>& v) {
  std::printf("[");
  for (size_t k = 0; k < v.size(); ++k) std::printf("%s%lld", k ? ", " : "", static_cast<long long>(v[k]));
  std::printf("]");
}

}  // namespace

int main() {
  Rng g(20260904);
  const Config k;
  std::printf("[\n");
  for (int t = 0; t < 120; ++t) {
    const Lp lp = make(g, k);
    std::printf(" {\"rows\": %d, \"cols\": %d, \"a\": [", lp.rows, lp.cols);
    for (int i = 0; i < lp.rows; ++i) {
      std::printf("%s", i ? ", " : "");
      list(lp.a[static_cast<size_t>(i)]);
    }
    std::printf("], \"b\": ");
    list(lp.b);
    std::printf(", \"c\": ");
    list(lp.c);
    std::printf(", \"upper\": ");
    list(lp.upper);
    std::printf(", \"x\": ");
    list(lp.x);
    std::printf(", \"objective\": %lld}%s\n", static_cast<long long>(lp.objective), t + 1 < 120 ? "," : "");
  }
  std::printf("]\n");
  return 0;
}
