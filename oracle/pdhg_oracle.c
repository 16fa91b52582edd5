/*
 * pdhg_oracle.c - CPU restatement of SANKHYA's restarted PDHG (see pdhg_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the timed CPU baseline.
 * Nothing in the product links it.
 *
 * Arithmetic is written in the reference's evaluation order so that, on the same input,
 * every deterministic output matches the reference bit for bit (double precision, no FMA
 * contraction: build with -ffp-contract=off).
 */
#define _POSIX_C_SOURCE 200809L
#include "pdhg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_MPS_INFINITY 1e30
#define OR_PRIMAL_FEAS 1e-7 /* tol::kPrimalFeasibility, include/sankhya/tolerances.hpp:19 */
#define OR_DUAL_FEAS 1e-7   /* tol::kDualFeasibility, tolerances.hpp:22 */
#define OR_DUALITY_GAP 1e-9 /* tol::kDualityGap, tolerances.hpp:108 */
#define OR_EVAL_INTERVAL 40 /* kEvaluationInterval, src/pdhg/pdhg.cpp:65 */
#define OR_RUIZ 10          /* kRuizIterations, pdhg.cpp:62 */
#define OR_POWER 30         /* kPowerIterations, pdhg.cpp:63 */

/* include/sankhya/types.hpp:89-96 */
int or_is_finite_bound(double v) { return !(isinf(v) || fabs(v) >= OR_MPS_INFINITY); }

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- src/la/sparse.cpp:232-269 ------------------------------------------------------- */

void or_csc_multiply(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                     const double* val, const double* x, double* y) {
  for (int32_t i = 0; i < m; ++i) y[i] = 0.0;
  for (int32_t j = 0; j < n; ++j) {
    const double xj = 1.0 * x[j];
    if (xj == 0.0) continue;
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) y[rowidx[k]] += val[k] * xj;
  }
}

/* Threads of the column gather, the one parallel loop of the reference's CPU engine
 * (src/la/sparse.cpp:244-259, OpenMP under option threads): every y[j] is written by one
 * thread, so the result is the same at any count. */
static int g_threads = 1;

void or_set_threads(int threads) { g_threads = threads > 0 ? threads : 1; }

static void csc_transpose_multiply_add(int32_t n, const int32_t* colptr, const int32_t* rowidx,
                                       const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int32_t j = 0; j < n; ++j) {
    double dot = 0.0;
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) dot += val[k] * x[rowidx[k]];
    y[j] += 1.0 * dot;
  }
}

void or_csc_transpose_multiply(int32_t m, int32_t n, const int32_t* colptr,
                               const int32_t* rowidx, const double* val, const double* x,
                               double* y) {
  (void)m;
  for (int32_t j = 0; j < n; ++j) y[j] = 0.0;
  csc_transpose_multiply_add(n, colptr, rowidx, val, x, y);
}

/* ---- src/la/scaling.cpp:14-159 ------------------------------------------------------- */

/* apply_pass (scaling.cpp:14-40). The reference rebuilds the matrix through triplets and
 * finalize(0.0); the pattern and order are unchanged by that (no duplicates, nothing below a
 * zero drop tolerance), so scaling the values in place gives the same numbers. */
static void apply_pass(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                       double* val, const double* row_pass, const double* col_pass, double* row,
                       double* col) {
  for (int32_t j = 0; j < n; ++j) {
    const double cj = col_pass[j];
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) val[k] = val[k] * row_pass[rowidx[k]] * cj;
  }
  for (int32_t i = 0; i < m; ++i) row[i] *= row_pass[i];
  for (int32_t j = 0; j < n; ++j) col[j] *= col_pass[j];
}

static double scale_bound(double bound, double multiplier) {
  return or_is_finite_bound(bound) ? bound * multiplier : bound;
}

void or_build_scaling(const or_model* model, const double* min_space_cost, int ruiz_iterations,
                      double* row, double* column, double* scaled_val, double* cost_hat,
                      double* col_lower_hat, double* col_upper_hat, double* row_lower_hat,
                      double* row_upper_hat, double* min_abs, double* max_abs) {
  const int32_t m = model->m, n = model->n;
  const int64_t nnz = model->nnz;
  double* rmax = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* cmax = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* rpass = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* cpass = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < m; ++i) row[i] = 1.0;
  for (int32_t j = 0; j < n; ++j) column[j] = 1.0;
  memcpy(scaled_val, model->val, sizeof(double) * (size_t)nnz);

  for (int pass = 0; pass <= ruiz_iterations; ++pass) {
    const int pock = (pass == ruiz_iterations); /* scaling.cpp:121-122 */
    for (int32_t i = 0; i < m; ++i) rmax[i] = 0.0;
    for (int32_t j = 0; j < n; ++j) cmax[j] = 0.0;
    for (int32_t j = 0; j < n; ++j) {
      for (int32_t k = model->colptr[j]; k < model->colptr[j + 1]; ++k) {
        const double mag = fabs(scaled_val[k]);
        const int32_t i = model->rowidx[k];
        if (pock) { /* scaling.cpp:82-89 */
          rmax[i] += mag;
          cmax[j] += mag;
        } else { /* scaling.cpp:50-59 */
          rmax[i] = dmax(rmax[i], mag);
          cmax[j] = dmax(cmax[j], mag);
        }
      }
    }
    for (int32_t i = 0; i < m; ++i) rpass[i] = rmax[i] > 0.0 ? 1.0 / sqrt(rmax[i]) : 1.0;
    for (int32_t j = 0; j < n; ++j) cpass[j] = cmax[j] > 0.0 ? 1.0 / sqrt(cmax[j]) : 1.0;
    apply_pass(m, n, model->colptr, model->rowidx, scaled_val, rpass, cpass, row, column);
  }

  for (int32_t j = 0; j < n; ++j) { /* scaling.cpp:130-136 */
    const double dc = column[j];
    cost_hat[j] = min_space_cost[j] * dc;
    col_lower_hat[j] = scale_bound(model->col_lower[j], 1.0 / dc);
    col_upper_hat[j] = scale_bound(model->col_upper[j], 1.0 / dc);
  }
  for (int32_t i = 0; i < m; ++i) { /* scaling.cpp:140-145 */
    const double dr = row[i];
    row_lower_hat[i] = scale_bound(model->row_lower[i], dr);
    row_upper_hat[i] = scale_bound(model->row_upper[i], dr);
  }
  {
    double mx = 0.0, mn = 0.0;
    int first = 1;
    for (int64_t k = 0; k < nnz; ++k) {
      const double mag = fabs(scaled_val[k]);
      mx = dmax(mx, mag);
      if (first || mag < mn) {
        mn = mag;
        first = 0;
      }
    }
    if (min_abs) *min_abs = mn;
    if (max_abs) *max_abs = mx;
  }
  free(rmax);
  free(cmax);
  free(rpass);
  free(cpass);
}

/* ---- std::mt19937 (C++ [rand.eng.mers], [rand.predef]) ------------------------------- */

void or_mt19937_seed(or_mt19937* g, uint32_t seed) {
  g->state[0] = seed;
  for (int i = 1; i < 624; ++i) {
    const uint32_t prev = g->state[i - 1];
    g->state[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
  }
  g->index = 624;
}

uint32_t or_mt19937_next(or_mt19937* g) {
  if (g->index >= 624) {
    for (int i = 0; i < 624; ++i) {
      const uint32_t y = (g->state[i] & 0x80000000u) | (g->state[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->state[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->state[i] = v;
    }
    g->index = 0;
  }
  uint32_t z = g->state[g->index++];
  z ^= (z >> 11);
  z ^= (z << 7) & 0x9d2c5680u;
  z ^= (z << 15) & 0xefc60000u;
  z ^= (z >> 18);
  return z;
}

/* libstdc++ generate_canonical<double, 53>(mt19937): two 32-bit draws, low word first, then
 * uniform_real_distribution returns canonical * (b - a) + a. */
double or_uniform_real(or_mt19937* g, double a, double b) {
  const double r = 4294967296.0;
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)or_mt19937_next(g) * tmp;
    tmp *= r;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret * (b - a) + a;
}

/* ---- src/la/scaling.cpp:161-194 ------------------------------------------------------ */

double or_spectral_norm(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                        const double* val, int iterations, uint32_t seed) {
  if (m == 0 || n == 0 || colptr[n] == 0) return 1.0;
  or_mt19937 g;
  or_mt19937_seed(&g, seed);
  double* v = (double*)malloc(sizeof(double) * (size_t)n);
  double* av = (double*)malloc(sizeof(double) * (size_t)m);
  for (int32_t j = 0; j < n; ++j) v[j] = or_uniform_real(&g, -1.0, 1.0);
  double norm = 0.0;
  for (int step = 0; step < iterations; ++step) {
    double length = 0.0;
    for (int32_t j = 0; j < n; ++j) length += v[j] * v[j];
    length = sqrt(length);
    if (length == 0.0) {
      free(v);
      free(av);
      return 1.0;
    }
    for (int32_t j = 0; j < n; ++j) v[j] /= length;
    or_csc_multiply(m, n, colptr, rowidx, val, v, av);
    or_csc_transpose_multiply(m, n, colptr, rowidx, val, av, v);
    double next = 0.0;
    for (int32_t j = 0; j < n; ++j) next += v[j] * v[j];
    norm = sqrt(sqrt(next));
  }
  free(v);
  free(av);
  return norm > 0.0 ? norm * 1.01 : 1.0;
}

/* ---- src/pdhg/pdhg_evaluate.cpp:19-139 ----------------------------------------------- */

void or_evaluate(const or_model* model, const double* min_cost, double bound_norm,
                 double cost_norm, const double* x, const double* y, double* activity,
                 double* reduced, or_residuals* r) {
  const int32_t m = model->m, n = model->n;
  memset(r, 0, sizeof(*r));
  for (int32_t i = 0; i < m; ++i) activity[i] = 0.0;
  if (m > 0) or_csc_multiply(m, n, model->colptr, model->rowidx, model->val, x, activity);
  double primal_violation = 0.0;
  for (int32_t i = 0; i < m; ++i) {
    const double a = activity[i];
    double violation = 0.0;
    if (or_is_finite_bound(model->row_lower[i])) violation = dmax(violation, model->row_lower[i] - a);
    if (or_is_finite_bound(model->row_upper[i])) violation = dmax(violation, a - model->row_upper[i]);
    primal_violation += violation * violation;
  }
  r->absolute_primal = sqrt(primal_violation);
  r->primal = r->absolute_primal / (1.0 + bound_norm);

  for (int32_t j = 0; j < n; ++j) reduced[j] = min_cost[j];
  if (m > 0) csc_transpose_multiply_add(n, model->colptr, model->rowidx, model->val, y, reduced);

  double dual_violation = 0.0, bound_contribution = 0.0;
  for (int32_t j = 0; j < n; ++j) {
    const double d = reduced[j];
    if (d > 0.0) {
      if (or_is_finite_bound(model->col_lower[j])) bound_contribution += d * model->col_lower[j];
      else dual_violation += d * d;
    } else if (d < 0.0) {
      if (or_is_finite_bound(model->col_upper[j])) bound_contribution += d * model->col_upper[j];
      else dual_violation += d * d;
    }
  }
  double support = 0.0;
  for (int32_t i = 0; i < m; ++i) {
    const double yi = y[i];
    if (yi > 0.0) {
      if (or_is_finite_bound(model->row_upper[i])) support += yi * model->row_upper[i];
      else dual_violation += yi * yi;
    } else if (yi < 0.0) {
      if (or_is_finite_bound(model->row_lower[i])) support += yi * model->row_lower[i];
      else dual_violation += yi * yi;
    }
  }
  r->absolute_dual = sqrt(dual_violation);
  r->dual = r->absolute_dual / (1.0 + cost_norm);

  for (int32_t i = 0; i < m; ++i) {
    if (model->row_lower[i] == model->row_upper[i]) continue;
    const double ls = or_is_finite_bound(model->row_lower[i]) ? activity[i] - model->row_lower[i] : INFINITY;
    const double us = or_is_finite_bound(model->row_upper[i]) ? model->row_upper[i] - activity[i] : INFINITY;
    r->complementarity = dmax(r->complementarity, fabs(y[i]) * dmin(ls, us));
  }
  for (int32_t j = 0; j < n; ++j) {
    if (model->col_lower[j] == model->col_upper[j]) continue;
    const double ls = or_is_finite_bound(model->col_lower[j]) ? x[j] - model->col_lower[j] : INFINITY;
    const double us = or_is_finite_bound(model->col_upper[j]) ? model->col_upper[j] - x[j] : INFINITY;
    r->complementarity = dmax(r->complementarity, fabs(reduced[j]) * dmin(ls, us));
  }
  double primal_objective = 0.0;
  for (int32_t j = 0; j < n; ++j) primal_objective += min_cost[j] * x[j];
  r->primal_objective = primal_objective;
  r->dual_objective = bound_contribution - support;
  const double absolute_gap = fabs(r->primal_objective - r->dual_objective);
  r->gap = absolute_gap / (1.0 + fabs(r->primal_objective) + fabs(r->dual_objective));
  r->gap_as_verified = absolute_gap / dmax(1.0, fabs(r->primal_objective));
}

/* pdhg_evaluate.hpp:48-69 */
static double worst(const or_residuals* r) { return dmax(dmax(r->primal, r->dual), r->gap); }
static int meets_request(const or_residuals* r, double t) {
  return r->primal <= t && r->dual <= t && r->gap <= t && r->absolute_primal <= OR_PRIMAL_FEAS;
}
static int meets_standard(const or_residuals* r) {
  return r->absolute_primal <= OR_PRIMAL_FEAS && r->absolute_dual <= OR_DUAL_FEAS &&
         r->gap_as_verified <= OR_DUALITY_GAP && r->complementarity <= 1e-6;
}

/* pdhg.cpp:68-72 */
static double project(double value, double lower, double upper) {
  if (or_is_finite_bound(lower) && value < lower) return lower;
  if (or_is_finite_bound(upper) && value > upper) return upper;
  return value;
}

static double euclid(const double* v, int64_t len) {
  double s = 0.0;
  for (int64_t k = 0; k < len; ++k) s += v[k] * v[k];
  return sqrt(s);
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

/* ---- src/pdhg/pdhg.cpp:76-568 -------------------------------------------------------- */

#define NEWV(len) ((double*)calloc((size_t)((len) > 0 ? (len) : 1), sizeof(double)))

int or_solve_pdhg(const or_model* model, const or_options* options, double* col_value,
                  double* row_dual, double* col_dual, or_result* result, or_tick* ticks,
                  int64_t max_ticks, int64_t* num_ticks) {
  const double t0 = now_seconds();
  const int32_t rows = model->m, cols = model->n;
  const size_t n = (size_t)cols, m = (size_t)rows;
  const double sense = model->maximize ? -1.0 : 1.0;
  int64_t tick_count = 0;
  memset(result, 0, sizeof(*result));

  /* validate() essentials (src/core/model.cpp:127-237) */
  for (int64_t k = 0; k < model->nnz; ++k) {
    if (!isfinite(model->val[k])) {
      result->status = OR_MODEL_ERROR;
      return result->status;
    }
  }
  for (int32_t j = 0; j < cols; ++j) {
    if (isnan(model->col_lower[j]) || isnan(model->col_upper[j]) ||
        model->col_lower[j] > model->col_upper[j] || isnan(model->cost[j])) {
      result->status = OR_MODEL_ERROR;
      return result->status;
    }
  }
  for (int32_t i = 0; i < rows; ++i) {
    if (isnan(model->row_lower[i]) || isnan(model->row_upper[i]) ||
        model->row_lower[i] > model->row_upper[i]) {
      result->status = OR_MODEL_ERROR;
      return result->status;
    }
  }

  /* Problem (pdhg.cpp:94-111) */
  double* cost = NEWV(n);
  for (size_t j = 0; j < n; ++j) cost[j] = sense * model->cost[j];
  const double cost_norm = euclid(cost, (int64_t)n);
  double bound_square = 0.0;
  for (size_t i = 0; i < m; ++i) {
    const double b = or_is_finite_bound(model->row_lower[i])
                         ? model->row_lower[i]
                         : (or_is_finite_bound(model->row_upper[i]) ? model->row_upper[i] : 0.0);
    bound_square += b * b;
  }
  const double bound_norm = sqrt(bound_square);

  /* Preconditioning (pdhg.cpp:114-117) */
  double* srow = NEWV(m);
  double* scol = NEWV(n);
  double* sval = NEWV(model->nnz);
  double* chat = NEWV(n);
  double* lhat = NEWV(n);
  double* uhat = NEWV(n);
  double* rlhat = NEWV(m);
  double* ruhat = NEWV(m);
  or_build_scaling(model, cost, OR_RUIZ, srow, scol, sval, chat, lhat, uhat, rlhat, ruhat, NULL,
                   NULL);
  const double spectral_norm = or_spectral_norm(rows, cols, model->colptr, model->rowidx, sval,
                                                OR_POWER, options->random_seed + 1u);
  result->spectral_norm = spectral_norm;

  const double tolerance = options->tolerance;
  const int64_t iteration_limit = options->iteration_limit < 0 ? 1000000 : options->iteration_limit;
  const int use_restarts = options->restart != 0;
  const int stop_at_request = options->stop_at_request != 0;
  const int has_time_limit = options->time_limit > 0.0 && isfinite(options->time_limit);

  double* x = NEWV(n);
  double* y = NEWV(m);
  for (size_t j = 0; j < n; ++j) x[j] = project(0.0, lhat[j], uhat[j]);
  double* x_next = NEWV(n);
  double* y_next = NEWV(m);
  double* ext = NEWV(n);
  double* at_y = NEWV(n);
  double* a_x = NEWV(m);
  double* x_sum = NEWV(n);
  double* y_sum = NEWV(m);
  int64_t averaged = 0;
  double* x_restart = NEWV(n);
  double* y_restart = NEWV(m);
  memcpy(x_restart, x, sizeof(double) * n);
  double* activity = NEWV(m);
  double* reduced = NEWV(n);
  double* cur_x = NEWV(n);
  double* cur_y = NEWV(m);
  double* avg_x = NEWV(n);
  double* avg_y = NEWV(m);
  double* best_x = NEWV(n);
  double* best_y = NEWV(m);
  double* dx = NEWV(n);
  double* adx = NEWV(m);
  double* dy = NEWV(m);
  memcpy(best_x, x, sizeof(double) * n); /* pdhg.cpp:190-191: best starts at the scaled x */

  double eta = spectral_norm > 0.0 ? 1.0 / spectral_norm : 1.0;
  double omega = 1.0;
  int64_t iteration = 0, restarts = 0, last_restart = 0;
  double restart_kkt = INFINITY;
  or_residuals best;
  memset(&best, 0, sizeof(best));
  best.primal = best.dual = best.gap = INFINITY;
  int converged = 0;
  int stop_status = OR_ITERATION_LIMIT;

  for (;;) {
    if (has_time_limit && now_seconds() - t0 >= options->time_limit) {
      stop_status = OR_TIME_LIMIT;
      break;
    }
    if (iteration >= iteration_limit) break;

    const double tau = eta / omega;
    const double sigma = eta * omega;

    /* Primal (pdhg.cpp:225-236) */
    for (size_t j = 0; j < n; ++j) at_y[j] = 0.0;
    if (rows > 0) or_csc_transpose_multiply(rows, cols, model->colptr, model->rowidx, sval, y, at_y);
    for (size_t j = 0; j < n; ++j) {
      const double gradient = chat[j] + at_y[j];
      x_next[j] = project(x[j] - tau * gradient, lhat[j], uhat[j]);
      ext[j] = 2.0 * x_next[j] - x[j];
    }
    /* Dual (pdhg.cpp:238-247) */
    if (rows > 0) or_csc_multiply(rows, cols, model->colptr, model->rowidx, sval, ext, a_x);
    for (size_t i = 0; i < m; ++i) {
      const double v = y[i] + sigma * a_x[i];
      y_next[i] = v - sigma * project(v / sigma, rlhat[i], ruhat[i]);
    }
    /* Movement (pdhg.cpp:252-260) */
    double movement = 0.0;
    for (size_t j = 0; j < n; ++j) {
      const double d = x_next[j] - x[j];
      movement += 0.5 * omega * d * d;
    }
    for (size_t i = 0; i < m; ++i) {
      const double d = y_next[i] - y[i];
      movement += 0.5 * d * d / omega;
    }
    /* Interaction (pdhg.cpp:262-277) */
    double interaction = 0.0;
    if (rows > 0) {
      for (size_t j = 0; j < n; ++j) dx[j] = x_next[j] - x[j];
      or_csc_multiply(rows, cols, model->colptr, model->rowidx, sval, dx, adx);
      for (size_t i = 0; i < m; ++i) interaction += (y_next[i] - y[i]) * adx[i];
      interaction = fabs(interaction);
    }
    /* Step rule (pdhg.cpp:293-335) */
    const int no_information = interaction <= 0.0;
    const double limit = no_information ? INFINITY : movement / interaction;
    const double exponent = (double)(iteration + 1 > 2 ? iteration + 1 : 2);
    const double shrink = 1.0 - pow(exponent, -0.3);
    const double grow = 1.0 + pow(exponent, -0.6);
    const double proposed = dmin(shrink * limit, grow * eta);
    if (eta <= limit) {
      double* t;
      t = x; x = x_next; x_next = t;
      t = y; y = y_next; y_next = t;
      for (size_t j = 0; j < n; ++j) x_sum[j] += x[j];
      for (size_t i = 0; i < m; ++i) y_sum[i] += y[i];
      ++averaged;
      ++iteration;
    }
    const double eta_ceiling = 1.0e3 / dmax(spectral_norm, 1e-12);
    if (!no_information) eta = clampd(proposed, 1e-12, eta_ceiling);

    /* Evaluation tick (pdhg.cpp:341-377) */
    if (iteration == 0) continue;
    if (iteration % OR_EVAL_INTERVAL != 0 && !no_information) continue;

    for (size_t j = 0; j < n; ++j) cur_x[j] = x[j] * scol[j];
    for (size_t i = 0; i < m; ++i) cur_y[i] = y[i] * srow[i];
    or_residuals current, average, better;
    or_evaluate(model, cost, bound_norm, cost_norm, cur_x, cur_y, activity, reduced, &current);
    int chose_average = 0;
    better = current;
    if (averaged > 0) {
      const double count = (double)averaged;
      for (size_t j = 0; j < n; ++j) avg_x[j] = (x_sum[j] / count) * scol[j];
      for (size_t i = 0; i < m; ++i) avg_y[i] = (y_sum[i] / count) * srow[i];
      or_evaluate(model, cost, bound_norm, cost_norm, avg_x, avg_y, activity, reduced, &average);
      if (worst(&average) < worst(&current)) {
        better = average;
        chose_average = 1;
      }
    }
    const double* chosen_x = chose_average ? avg_x : cur_x;
    const double* chosen_y = chose_average ? avg_y : cur_y;
    if (worst(&better) < worst(&best)) { /* pdhg.cpp:379-384 */
      best = better;
      memcpy(best_x, chosen_x, sizeof(double) * n);
      memcpy(best_y, chosen_y, sizeof(double) * m);
    }
    int restarted = 0;
    const int stop_here =
        options->stop_rule == 1
            ? worst(&better) <= tolerance
            : meets_request(&better, tolerance) && (stop_at_request || meets_standard(&better));
    if (stop_here) { /* pdhg.cpp:424-436 */
      best = better;
      memcpy(best_x, chosen_x, sizeof(double) * n);
      memcpy(best_y, chosen_y, sizeof(double) * m);
      converged = 1;
    } else if (use_restarts) { /* pdhg.cpp:438-483 */
      const double kkt = worst(&better);
      const int64_t since = iteration - last_restart;
      const int sufficient = kkt <= 0.2 * restart_kkt;
      int64_t art = (int64_t)(0.36 * (double)iteration);
      if (art < OR_EVAL_INTERVAL) art = OR_EVAL_INTERVAL;
      const int artificial = since >= art;
      if (sufficient || artificial) {
        for (size_t j = 0; j < n; ++j) dx[j] = x[j] - x_restart[j];
        for (size_t i = 0; i < m; ++i) dy[i] = y[i] - y_restart[i];
        const double dx_norm = euclid(dx, (int64_t)n);
        const double dy_norm = euclid(dy, (int64_t)m);
        if (dx_norm > 1e-12 && dy_norm > 1e-12) {
          const double theta = 0.5;
          omega = exp(theta * log(dy_norm / dx_norm) + (1.0 - theta) * log(omega));
          omega = clampd(omega, 1e-6, 1e6);
        }
        memset(x_sum, 0, sizeof(double) * n);
        memset(y_sum, 0, sizeof(double) * m);
        averaged = 0;
        memcpy(x_restart, x, sizeof(double) * n);
        memcpy(y_restart, y, sizeof(double) * m);
        restart_kkt = kkt;
        last_restart = iteration;
        ++restarts;
        restarted = 1;
      }
    }
    if (ticks && tick_count < max_ticks) {
      or_tick* tk = &ticks[tick_count];
      tk->iteration = iteration;
      tk->objective = sense * better.primal_objective + model->objective_offset;
      tk->primal = better.primal;
      tk->dual = better.dual;
      tk->gap = better.gap;
      tk->restarted = restarted;
      tk->chose_average = chose_average;
    }
    ++tick_count;
    if (converged) break;
  }

  /* Report (pdhg.cpp:486-567) */
  for (size_t j = 0; j < n; ++j) col_value[j] = best_x[j];
  or_residuals final_r;
  or_evaluate(model, cost, bound_norm, cost_norm, best_x, best_y, activity, reduced, &final_r);
  for (size_t j = 0; j < n; ++j) col_dual[j] = sense * reduced[j];
  for (size_t i = 0; i < m; ++i) row_dual[i] = sense * (-best_y[i]);
  result->iterations = iteration;
  result->restarts = restarts;
  result->final_residuals = final_r;
  /* A relative-KKT stop claims nothing unless the reference's request holds too (the engine's
   * report does the same). */
  if (converged && options->stop_rule == 1 && !meets_request(&final_r, tolerance)) converged = 0;
  const int verifiable = converged && meets_standard(&final_r);
  if (verifiable) result->status = OR_OPTIMAL;
  else if (converged) result->status = OR_FEASIBLE;
  else result->status = (stop_status == OR_TIME_LIMIT || stop_status == OR_INTERRUPTED) ? stop_status : OR_ITERATION_LIMIT;
  if (verifiable) result->dual_bound = sense * final_r.dual_objective + model->objective_offset;
  else result->dual_bound = model->maximize ? INFINITY : -INFINITY;
  result->solve_seconds = now_seconds() - t0;
  {
    const double db = result->dual_bound;
    const int32_t st = result->status;
    or_recompute_quality(model, col_value, row_dual, col_dual, db, activity, result);
    result->status = st;
  }
  if (num_ticks) *num_ticks = tick_count;

  free(cost); free(srow); free(scol); free(sval); free(chat); free(lhat); free(uhat);
  free(rlhat); free(ruhat); free(x); free(y); free(x_next); free(y_next); free(ext);
  free(at_y); free(a_x); free(x_sum); free(y_sum); free(x_restart); free(y_restart);
  free(activity); free(reduced); free(cur_x); free(cur_y); free(avg_x); free(avg_y);
  free(best_x); free(best_y); free(dx); free(adx); free(dy);
  return result->status;
}

/* ---- src/core/model.cpp:247-441 (LP: no integrality, no Hessian) --------------------- */

static double nearest_bound_distance(double value, double lo, double hi) {
  double d = INFINITY;
  if (or_is_finite_bound(lo)) d = dmin(d, fabs(value - lo));
  if (or_is_finite_bound(hi)) d = dmin(d, fabs(value - hi));
  return d;
}

void or_recompute_quality(const or_model* model, const double* col_value,
                          const double* row_dual, const double* col_dual, double dual_bound,
                          double* row_activity, or_result* q) {
  const int32_t n = model->n, m = model->m;
  const double sense = model->maximize ? -1.0 : 1.0;
  if (m > 0) or_csc_multiply(m, n, model->colptr, model->rowidx, model->val, col_value, row_activity);
  double pinf = 0.0, pinf_s = 0.0;
  double* row_scale = NEWV(m);
  for (int32_t i = 0; i < m; ++i) row_scale[i] = 1.0;
  for (int32_t j = 0; j < n; ++j) {
    const double x = col_value[j];
    if (x == 0.0) continue;
    for (int32_t k = model->colptr[j]; k < model->colptr[j + 1]; ++k) {
      const int32_t r = model->rowidx[k];
      row_scale[r] = dmax(row_scale[r], fabs(model->val[k] * x));
    }
  }
#define RECORD(viol, scale)                                    \
  do {                                                         \
    const double v_ = (viol);                                  \
    if (v_ > 0.0) {                                            \
      pinf = dmax(pinf, v_);                                   \
      pinf_s = dmax(pinf_s, v_ / dmax(1.0, (scale)));          \
    }                                                          \
  } while (0)
  for (int32_t j = 0; j < n; ++j) {
    const double x = col_value[j];
    if (or_is_finite_bound(model->col_lower[j])) RECORD(model->col_lower[j] - x, fabs(x));
    if (or_is_finite_bound(model->col_upper[j])) RECORD(x - model->col_upper[j], fabs(x));
  }
  for (int32_t i = 0; i < m; ++i) {
    const double a = row_activity[i];
    if (or_is_finite_bound(model->row_lower[i])) RECORD(model->row_lower[i] - a, row_scale[i]);
    if (or_is_finite_bound(model->row_upper[i])) RECORD(a - model->row_upper[i], row_scale[i]);
  }
#undef RECORD
  q->primal_infeasibility = pinf;
  q->primal_infeasibility_scaled = pinf_s;
  double objective = model->objective_offset;
  for (int32_t j = 0; j < n; ++j) objective += model->cost[j] * col_value[j];
  q->objective = objective;

  double dinf = 0.0, dinf_s = 0.0, compl = 0.0;
  double dual_norm = 0.0;
  for (int32_t i = 0; i < m; ++i) dual_norm = dmax(dual_norm, fabs(row_dual[i]));
#define RECORD_DUAL(viol, scale)                  \
  do {                                            \
    const double v_ = (viol);                     \
    if (v_ > 0.0) {                               \
      dinf = dmax(dinf, v_);                      \
      dinf_s = dmax(dinf_s, v_ / (scale));        \
    }                                             \
  } while (0)
#define RECORD_COMPL(mult, slack, mscale, pscale)                                  \
  do {                                                                             \
    const double mu_ = (mult), sl_ = (slack);                                      \
    if (mu_ != 0.0 && or_is_finite_bound(sl_)) {                                   \
      const double pr_ = mu_ * sl_;                                                \
      compl = dmax(compl, pr_);                                                    \
      dinf_s = dmax(dinf_s, pr_ / dmax(1.0, (mscale) * (pscale)));                 \
    }                                                                              \
  } while (0)
  for (int32_t j = 0; j < n; ++j) {
    double implied = model->cost[j];
    double terms = fabs(model->cost[j]);
    for (int32_t k = model->colptr[j]; k < model->colptr[j + 1]; ++k) {
      const double term = model->val[k] * row_dual[model->rowidx[k]];
      implied -= term;
      terms = dmax(terms, fabs(term));
    }
    RECORD_DUAL(fabs(implied - col_dual[j]), dmax(1.0, terms));
  }
  for (int32_t j = 0; j < n; ++j) {
    const double d = sense * col_dual[j];
    const double x = col_value[j];
    const double lo = model->col_lower[j], hi = model->col_upper[j];
    if (lo == hi) continue;
    double scale = fabs(model->cost[j]);
    for (int32_t k = model->colptr[j]; k < model->colptr[j + 1]; ++k)
      scale = dmax(scale, fabs(model->val[k] * row_dual[model->rowidx[k]]));
    scale = dmax(1.0, scale);
    if (d > 0.0 && !or_is_finite_bound(lo)) RECORD_DUAL(d, scale);
    if (d < 0.0 && !or_is_finite_bound(hi)) RECORD_DUAL(-d, scale);
    RECORD_COMPL(fabs(d), nearest_bound_distance(x, lo, hi), scale, dmax(1.0, fabs(x)));
  }
  for (int32_t i = 0; i < m; ++i) {
    const double yv = sense * row_dual[i];
    const double a = row_activity[i];
    const double lo = model->row_lower[i], hi = model->row_upper[i];
    if (lo == hi) continue;
    const double price_scale = dmax(1.0, dual_norm);
    if (yv > 0.0 && !or_is_finite_bound(lo)) RECORD_DUAL(yv, price_scale);
    if (yv < 0.0 && !or_is_finite_bound(hi)) RECORD_DUAL(-yv, price_scale);
    RECORD_COMPL(fabs(yv), nearest_bound_distance(a, lo, hi), price_scale, row_scale[i]);
  }
#undef RECORD_DUAL
#undef RECORD_COMPL
  q->dual_infeasibility = dinf;
  q->dual_infeasibility_scaled = dinf_s;
  q->complementarity_violation = compl;
  q->absolute_gap = fabs(objective - dual_bound);
  q->relative_gap = q->absolute_gap / dmax(1.0, fabs(objective));
  free(row_scale);
}
