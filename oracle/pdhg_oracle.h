/*
 * pdhg_oracle.h - CPU restatement of SANKHYA's restarted-PDHG LP engine, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP engine is compared
 * against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * only as the checker or as the timed CPU baseline; the product never links or calls it.
 *
 * Each function follows one reference function operation for operation (summation order
 * included) so that its outputs can be compared bit-for-bit where the reference's are
 * deterministic:
 *
 *   or_csc_multiply              src/la/sparse.cpp:232-242,261-264   (column scatter, skips x_j==0)
 *   or_csc_transpose_multiply    src/la/sparse.cpp:244-259,266-269   (column gather)
 *   or_build_scaling             src/la/scaling.cpp:14-159           (10 Ruiz + 1 Pock-Chambolle)
 *   or_spectral_norm             src/la/scaling.cpp:161-194          (mt19937 + libstdc++ canonical)
 *   or_evaluate                  src/pdhg/pdhg_evaluate.cpp:19-139
 *   or_solve_pdhg                src/pdhg/pdhg.cpp:76-568
 *   or_recompute_quality         src/core/model.cpp:247-441          (LP subset of the fields)
 *
 * Pinning (what checks each function beyond retyping it) is listed in tests/test_oracle.py.
 */
#ifndef SANKHYA_PDHG_ORACLE_H
#define SANKHYA_PDHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the order of sankhya::SolveStatus (include/sankhya/model.hpp:64-96). */
enum {
  OR_NOT_SOLVED = 0,
  OR_OPTIMAL = 1,
  OR_INFEASIBLE = 2,
  OR_UNBOUNDED = 3,
  OR_INFEASIBLE_OR_UNBOUNDED = 4,
  OR_FEASIBLE = 5,
  OR_ITERATION_LIMIT = 6,
  OR_TIME_LIMIT = 7,
  OR_NODE_LIMIT = 8,
  OR_NUMERICAL_ERROR = 9,
  OR_MODEL_ERROR = 10,
  OR_INTERRUPTED = 11
};

/* A frozen LP in the reference's CSC layout (rows ascending inside each column). */
typedef struct {
  int32_t m, n;
  int64_t nnz;
  const int32_t* colptr; /* n+1 */
  const int32_t* rowidx; /* nnz */
  const double* val;     /* nnz */
  const double* cost;    /* n, in the model's own sense */
  const double* col_lower, *col_upper; /* n */
  const double* row_lower, *row_upper; /* m */
  double objective_offset;
  int32_t maximize; /* 0 = minimise, 1 = maximise */
} or_model;

typedef struct {
  double tolerance;        /* pdhg_tolerance */
  int64_t iteration_limit; /* <0 means the engine's own ceiling of 1e6 */
  int32_t restart;         /* pdhg_restart */
  int32_t stop_at_request; /* pdhg_stop_at_request */
  uint32_t random_seed;    /* the power iteration uses random_seed + 1 */
  double time_limit;       /* seconds, <=0 or inf: none */
  int32_t stop_rule;       /* 0: the reference's test; 1: the engine's pdhg_stop_rule=relative_kkt
                              (max of relative primal, dual, gap <= tolerance), for comparisons
                              with the engine run that way */
} or_options;

/* Residuals as in src/pdhg/pdhg_evaluate.hpp:19-46. */
typedef struct {
  double primal, dual, gap, primal_objective, dual_objective;
  double absolute_primal, absolute_dual, gap_as_verified, complementarity;
} or_residuals;

typedef struct {
  int32_t status;
  int64_t iterations;
  int64_t restarts;
  double objective;  /* recompute_quality's objective, in the model's sense, with offset */
  double dual_bound;
  double solve_seconds;
  double spectral_norm;
  or_residuals final_residuals;
  /* Quality fields of Solution::recompute_quality (LP subset). */
  double primal_infeasibility, primal_infeasibility_scaled;
  double dual_infeasibility, dual_infeasibility_scaled, complementarity_violation;
  double absolute_gap, relative_gap;
} or_result;

/* Tick log record: one per evaluation tick (pdhg.cpp:341-400). */
typedef struct {
  int64_t iteration;
  double objective; /* sense * better.primal_objective + offset */
  double primal, dual, gap;
  int32_t restarted;
  int32_t chose_average;
} or_tick;

int or_is_finite_bound(double v);

/* Threads for the A^T y gather (the reference's `threads` option, src/la/sparse.cpp:248-250);
 * 1 by default; results do not depend on it. */
void or_set_threads(int threads);

void or_csc_multiply(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                     const double* val, const double* x, double* y);
void or_csc_transpose_multiply(int32_t m, int32_t n, const int32_t* colptr,
                               const int32_t* rowidx, const double* val, const double* x,
                               double* y);

/* Ruiz(ruiz_iterations) + Pock-Chambolle. Outputs: row (m), column (n), the scaled CSC values
 * (nnz, same pattern), and the scaled cost/bounds. min_space_cost is sense*c. */
void or_build_scaling(const or_model* model, const double* min_space_cost, int ruiz_iterations,
                      double* row, double* column, double* scaled_val, double* cost_hat,
                      double* col_lower_hat, double* col_upper_hat, double* row_lower_hat,
                      double* row_upper_hat, double* min_abs, double* max_abs);

double or_spectral_norm(int32_t m, int32_t n, const int32_t* colptr, const int32_t* rowidx,
                        const double* val, int iterations, uint32_t seed);

/* std::mt19937 and std::uniform_real_distribution<double>(a,b) as libstdc++ implements them. */
typedef struct {
  uint32_t state[624];
  int index;
} or_mt19937;
void or_mt19937_seed(or_mt19937* g, uint32_t seed);
uint32_t or_mt19937_next(or_mt19937* g);
double or_uniform_real(or_mt19937* g, double a, double b);

/* evaluate(): x (n), y (m) unscaled; activity (m) and reduced (n) are outputs. cost is the
 * minimise-space cost; bound_norm/cost_norm as in pdhg.cpp:95-111. */
void or_evaluate(const or_model* model, const double* min_cost, double bound_norm,
                 double cost_norm, const double* x, const double* y, double* activity,
                 double* reduced, or_residuals* out);

/* The whole engine. col_value/col_dual (n) and row_dual (m) receive the Solution vectors.
 * ticks (may be NULL) receives up to max_ticks records; *num_ticks the count written. */
int or_solve_pdhg(const or_model* model, const or_options* options, double* col_value,
                  double* row_dual, double* col_dual, or_result* result, or_tick* ticks,
                  int64_t max_ticks, int64_t* num_ticks);

/* Solution::recompute_quality for an LP (no integrality, no Hessian). row_activity (m) out. */
void or_recompute_quality(const or_model* model, const double* col_value,
                          const double* row_dual, const double* col_dual, double dual_bound,
                          double* row_activity, or_result* quality);

#ifdef __cplusplus
}
#endif

#endif
