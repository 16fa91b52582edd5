/*
 * sankhya_pdhg.h - C ABI of the MI355X PDHG engine for foreign callers (Python ctypes, tests,
 * the benchmark).
 *
 * It plays the role of the reference's C API for this one path: sankhya_solve
 * (include/sankhya/sankhya.h:248, src/api/c_api.cpp:563) over a model built from arrays, with
 * options given as "name=value" text the way sankhya_options_set_* take them
 * (include/sankhya/sankhya.h:228-233). Plain C types only; caller-owned buffers.
 */
#ifndef SANKHYA_PDHG_H
#define SANKHYA_PDHG_H

#include <stdint.h>

#include "pdhg_batch.h"
#include "pdhg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sankhya_model sankhya_model;

/* A model from CSC arrays (rows ascending within each column); costs in the model's own
 * sense; maximize = 0 or 1. Returns NULL on a malformed pattern. */
sankhya_model* sankhya_model_new(int32_t m, int32_t n, int64_t nnz, const int32_t* colptr,
                                 const int32_t* rowidx, const double* val, const double* cost,
                                 const double* col_lower, const double* col_upper,
                                 const double* row_lower, const double* row_upper,
                                 double objective_offset, int32_t maximize);

/* A deterministic synthetic model: kind 0 uniform-equality, 1 zipf, 2 blend (the
 * constructions are documented with the generator). */
sankhya_model* sankhya_model_generate(int32_t kind, int32_t m, int32_t n, int32_t per_col,
                                      double param, uint64_t seed);

void sankhya_model_dims(const sankhya_model* model, int32_t* m, int32_t* n, int64_t* nnz);

/* Copies the model out (any pointer may be NULL). */
void sankhya_model_export(const sankhya_model* model, int32_t* colptr, int32_t* rowidx,
                          double* val, double* cost, double* col_lower, double* col_upper,
                          double* row_lower, double* row_upper, double* objective_offset,
                          int32_t* maximize);

void sankhya_model_free(sankhya_model* model);

/* Model and row/column names for solution files (the reference Model's name, source_path,
 * col_names, row_names; include/sankhya/model.hpp). Any pointer may be NULL (left unchanged);
 * col_names holds n entries and row_names m, a NULL entry meaning the default C<j> / R<i>. */
int sankhya_model_set_names(sankhya_model* model, const char* name, const char* source_path,
                            const char* const* col_names, const char* const* row_names);

/* The names back: the model's name (kind 0), its source path (1), column `index`'s name (2) or
 * row `index`'s name (3); NULL when there is none or the index is out of range. The pointer
 * stays valid until the model is changed or freed. */
const char* sankhya_model_name_of(const sankhya_model* model, int32_t kind, int32_t index);

/* 1 for an integer column (MPS MARKER, BV, LI or UI), 0 for a continuous one, -1 when `col` is
 * out of range. The PDHG engine solves the relaxation either way. */
int32_t sankhya_model_col_is_integer(const sankhya_model* model, int32_t col);

typedef struct {
  int32_t status; /* sankhya::SolveStatus */
  int64_t iterations, attempts, restarts, ticks, device_bytes;
  int64_t fp32_iterations; /* mixed precision: iterations run before the FP64 continuation */
  int32_t schedule;        /* batched engine: -1 cooperative, else the persistent geometry the
                              tuner chose (0: 256 threads, 1: 512, 2: 1024); -1 for one LP */
  double objective, dual_bound, solve_seconds, setup_seconds, loop_seconds, spectral_norm;
  double primal_infeasibility, dual_infeasibility, complementarity_violation, relative_gap;
  double final_primal, final_dual, final_gap;
  double final_absolute_primal; /* ||primal violation||_2 in the model's units at the report */
  char algorithm[32];
  char message[1024];
} sankhya_result;

/* Solve with the MI355X engine. options: "name=value" pairs separated by ';' or ','
 * (NULL or "" for defaults; log_to_console defaults to false here). Any output pointer may
 * be NULL. Returns 0, or -1 when an option is rejected (message says which). */
int sankhya_pdhg_solve(const sankhya_model* model, const char* options, double* col_value,
                       double* row_dual, double* col_dual, double* row_activity,
                       sankhya_result* result);

/* Solve nlp independent LPs with the batched engine (include/pdhg_batch.h); same options as
 * sankhya_pdhg_solve plus pdhg_batch_tune / pdhg_batch_tune_attempts. Outputs concatenated in
 * model order: col_value and col_dual hold sum(n) values, row_dual and row_activity sum(m);
 * results holds nlp records. Any output pointer may be NULL. Returns 0, or -1 when an option is
 * rejected (results[0].message says which). */
int sankhya_pdhg_solve_batch(const sankhya_model* const* models, int32_t nlp, const char* options,
                             double* col_value, double* row_dual, double* col_dual,
                             double* row_activity, sankhya_result* results);

/* A solved batch kept on the device for vector-Jacobian products (the backward pass of a
 * differentiable LP layer). sankhya_batch_solve solves like sankhya_pdhg_solve_batch and returns
 * the session (NULL when an option is rejected; results[0].message says which). */
typedef struct sankhya_batch sankhya_batch;
sankhya_batch* sankhya_batch_solve(const sankhya_model* const* models, int32_t nlp,
                                   const char* options, double* col_value, double* row_dual,
                                   double* col_dual, double* row_activity,
                                   sankhya_result* results);

/* dL/dc (sum n) and dL/d(active row bound) (sum m) of every LP for upstream dL/d(col_value)
 * (sum n) and dL/d(row_dual) (sum m), in the models' own sense and sign conventions; inactive
 * rows and columns at a bound get 0. Implicit differentiation of the PDHG fixed point on the
 * active set, FP64 CGLS on the device (include/pdhg_batch.h, pdhg_batch_adjoint): tol is the
 * relative normal-equation residual at which it stops. iterations and residuals: 2 per LP (may
 * be NULL). Returns 0 or a PDHG_E_* code. */
int sankhya_batch_adjoint(const sankhya_batch* session, const double* grad_col_value,
                          const double* grad_row_dual, double tol, int max_iterations,
                          double* grad_cost, double* grad_row_bound, int32_t* iterations,
                          double* residuals);

/* Test hook: the active set of the last sankhya_batch_adjoint call, in model order: mask_c
 * (sum n) 1.0 on free columns, mask_r (sum m) 1.0 on active rows, 0.0 elsewhere. */
int sankhya_batch_adjoint_masks(const sankhya_batch* session, double* mask_c, double* mask_r);

void sankhya_batch_free(sankhya_batch* session);

/* Write the solution file, the engine's certificate output (replaces io::write_solution,
 * include/sankhya/io.hpp:65-72, format of src/io/writer.cpp:121-284), from a result and its
 * vectors as sankhya_pdhg_solve returned them (NULL vectors are written as zeros). options:
 * as for sankhya_pdhg_solve; mip_relative_gap / mip_absolute_gap go into the header. Returns
 * 0, or -1 with the reason in error (error_len bytes, may be NULL). */
int sankhya_write_solution(const sankhya_model* model, const char* options, const char* path,
                           const sankhya_result* result, const double* col_value,
                           const double* row_dual, const double* col_dual,
                           const double* row_activity, char* error, int32_t error_len);

/* The host preconditioner on its own (src/la/scaling.cpp:111-159): Dr (m), Dc (n), the scaled
 * CSC values (nnz), scaled cost/bounds. min_cost is sense * c. threads: host threads for the
 * passes (same result bit for bit; models under 200,000 nonzeros use one). */
int sankhya_pdhg_scaling(const sankhya_model* model, int ruiz_iterations, int threads, double* row,
                         double* column, double* scaled_val, double* cost_hat,
                         double* col_lower_hat, double* col_upper_hat, double* row_lower_hat,
                         double* row_upper_hat);

/* The power iteration's start vector (n) the engine draws for this model: std::mt19937 seeded
 * random_seed + 1, uniform on [-1, 1) (src/la/scaling.cpp:168-171). */
int sankhya_pdhg_power_start(const sankhya_model* model, int64_t random_seed, double* v0);

/* The CSR of the scaled matrix Ahat as the host preparation builds it (host preconditioner,
 * then the counting-sort transpose of src/la/sparse.cpp:271-302, CsrView::build :315-322):
 * rowptr (m + 1), colidx and values (nnz), columns ascending within each row. threads: host
 * threads for the scaling and the transpose (the same arrays for any count). The batched
 * engine uploads this; the single-LP engine uploads A's own CSR and scales on the device. */
int sankhya_pdhg_csr(const sankhya_model* model, int threads, int32_t* rowptr, int32_t* colidx,
                     double* values);

/* Opens a device handle holding this model exactly as the engine uploads it (both CSR layouts,
 * preconditioned on the device by pdhg_dev_upload_unscaled), for the pdhg_dev_* test hooks. Also returns the unscaled min-space cost norm
 * and bound norm through the handle. */
int sankhya_pdhg_device_open(const sankhya_model* model, int device, int precision,
                             pdhg_dev** out);

/* Opens a batch handle holding these models exactly as the batched engine uploads them (the
 * unscaled problems, preconditioned on the device by pdhg_batch_upload_unscaled), for the
 * pdhg_batch_* test hooks. */
int sankhya_pdhg_batch_open(const sankhya_model* const* models, int32_t nlp, int device,
                            int precision, pdhg_batch** out);

/* Solution::recompute_quality of the engine (the quality fields solve() reports; reference
 * src/core/model.cpp:247-441) for the point (col_value n, row_dual m, col_dual n) with the given
 * dual bound. row_activity (m) out; out[9] = objective, primal_infeasibility,
 * primal_infeasibility_scaled, dual_infeasibility, dual_infeasibility_scaled,
 * complementarity_violation, integrality_violation, absolute_gap, relative_gap. Host only. */
int sankhya_pdhg_quality(const sankhya_model* model, const double* col_value, const double* row_dual,
                         const double* col_dual, double dual_bound, double* row_activity,
                         double* out);

/* The same quality fields by the threaded form solve() uses after a device solve (the passes
 * over the nonzeros on `threads` host threads, rows through the CSR of A): bit for bit the values
 * of sankhya_pdhg_quality. No reference counterpart (a test hook). */
int sankhya_pdhg_quality_threaded(const sankhya_model* model, const double* col_value,
                                  const double* row_dual, const double* col_dual, double dual_bound,
                                  int threads, double* row_activity, double* out);

/* The engine's VRAM check on the host, as solve() makes it before allocating anything
 * (src/core/solve.cpp:751-771): the device-memory estimate of this model for the precision
 * (pdhg_dev_estimate_bytes) plus the reserve (pdhg_dev_vram_reserve) against free_bytes.
 * Returns 1 when it fits, 0 when it does not, -1 on bad arguments; required and reserve (may be
 * NULL) receive the two figures. Host arithmetic only: no device is needed. */
int sankhya_pdhg_memory_check(const sankhya_model* model, int precision, int64_t free_bytes,
                              int64_t total_bytes, int64_t* required, int64_t* reserve);

#ifdef __cplusplus
}
#endif

#endif
