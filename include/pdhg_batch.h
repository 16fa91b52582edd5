/*
 * pdhg_batch.h - the C ABI of the MI355X batched restarted-PDHG engine: many independent LPs
 * of uneven size solved together, all of them advanced by one captured HIP graph per chunk.
 *
 * The reference has no batched engine: it solves one LP per call of solve_pdhg
 * (include/sankhya/pdhg.hpp:15-16) and its CUDA path launches several kernels plus a blocking
 * copy per attempt (src/gpu/pdhg_gpu.cu:542-744). Each LP here follows the same algorithm as
 * the single-LP engine (include/pdhg_hip.h), decision for decision; what changes is the
 * schedule.
 *
 * Two schedules. Cooperative (the default): every LP's work items share one grid per
 * product, so an attempt of all cooperative LPs is a few launches with per-LP control blocks
 * and per-LP tail workgroups (the block-diagonal form of the single-LP engine). Persistent: one
 * workgroup per LP runs whole chunks of attempts and ticks with workgroup barriers only, in one
 * of three geometries (256, 512 or 1024 threads; stream tiles of 8x that many nonzeros); it
 * wins on LPs small enough for one CU. pdhg_batch_tune assigns each LP by timing; the
 * cooperative branch and each persistent bin run as concurrent branches of the captured graph.
 *
 * Conventions as in pdhg_hip.h: plain C types, 0 = PDHG_OK or a PDHG_E_* code, host pointers
 * copied, no CPU fallback. Arrays are concatenated LP after LP; indices inside an LP's arrays
 * are local to that LP.
 */
#ifndef SANKHYA_PDHG_BATCH_H
#define SANKHYA_PDHG_BATCH_H

#include <stdint.h>

#include "pdhg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdhg_batch pdhg_batch;

/* precision: PDHG_PRECISION_FP64, or PDHG_PRECISION_MIXED (FP32 matrix values and iterate
 * deltas around FP64 restart points; FP64 residuals, restarts and step control). */
int pdhg_batch_create(int device, int precision, pdhg_batch** out);
void pdhg_batch_destroy(pdhg_batch* h);
const char* pdhg_batch_last_error(const pdhg_batch* h);
int64_t pdhg_batch_bytes(const pdhg_batch* h);

/* nlp scaled LPs, as pdhg_dev_upload takes one (there: the CSR of Ahat and of Ahat^T, the
 * scaled vectors, Dc/Dr, the unscaled cost and bounds). m[k], n[k], nnz[k] per LP.
 * a_rowptr holds m[k]+1 entries per LP (values 0..nnz[k]), a_col/a_val nnz[k] per LP (column
 * indices 0..n[k]-1); at_* the same for Ahat^T. n-vectors hold n[k] entries per LP, m-vectors
 * m[k]; bound_norm and cost_norm one per LP. */
int pdhg_batch_upload(pdhg_batch* h, int32_t nlp, const int32_t* m, const int32_t* n,
                      const int64_t* nnz, const int32_t* a_rowptr, const int32_t* a_col,
                      const double* a_val, const int32_t* at_rowptr, const int32_t* at_col,
                      const double* at_val, const double* cost_hat, const double* lower_hat,
                      const double* upper_hat, const double* row_lower_hat,
                      const double* row_upper_hat, const double* col_scale,
                      const double* row_scale, const double* min_cost, const double* col_lower,
                      const double* col_upper, const double* row_lower, const double* row_upper,
                      const double* bound_norm, const double* cost_norm);

/* The same upload from each LP's UNSCALED problem, preconditioned on the device as
 * pdhg_dev_upload_unscaled does for one LP: `ruiz_iterations` Ruiz rounds and one Pock-Chambolle
 * round per LP (src/la/scaling.cpp:14-159), bit-identical to the host passes, then the scaled
 * cost and bounds. a_val / at_val hold A's own values; min_cost is sense * c per LP; the norms
 * are of the unscaled problems (src/pdhg/pdhg.cpp:94-111). */
int pdhg_batch_upload_unscaled(pdhg_batch* h, int32_t nlp, const int32_t* m, const int32_t* n,
                               const int64_t* nnz, const int32_t* a_rowptr, const int32_t* a_col,
                               const double* a_val, const int32_t* at_rowptr, const int32_t* at_col,
                               const double* at_val, const double* min_cost,
                               const double* col_lower, const double* col_upper,
                               const double* row_lower, const double* row_upper,
                               const double* bound_norm, const double* cost_norm,
                               int ruiz_iterations);

/* Test hook: the batch's preconditioning results, concatenated LP after LP: Dr (sum m), Dc
 * (sum n), the scaled values in CSC order (sum nnz) and the scaled cost and bounds; any pointer
 * may be null. */
int pdhg_batch_scaling(pdhg_batch* h, double* row, double* column, double* csc_values,
                       double* cost_hat, double* col_lower_hat, double* col_upper_hat,
                       double* row_lower_hat, double* row_upper_hat);

/* Power iteration per LP (src/la/scaling.cpp:161-194): v0 holds each LP's start vector
 * (n[k] entries per LP); norms receives nlp values, each already rounded up by 1.01. */
int pdhg_batch_spectral_norms(pdhg_batch* h, const double* v0, int iterations, double* norms);

/* The schedule tuner. Times `attempts` attempts of the smallest LP alone in the cooperative
 * schedule (HIP events: the cooperative floor, what an attempt costs once the other LPs have
 * stopped), then of every LP alone in a persistent workgroup at each geometry (device clock),
 * from the start point each time. An LP moves to its fastest persistent geometry only when that
 * beats the cooperative floor. geometry_out (may be NULL) receives -1 for cooperative, else
 * the geometry index (0: 256 threads, 1: 512, 2: 1024). Leaves the engine unstarted. Without
 * a call, every LP is cooperative. */
int pdhg_batch_tune(pdhg_batch* h, const pdhg_params* p, const double* spectral_norms,
                    int attempts, int32_t* geometry_out /* nlp, may be NULL */);

/* Starts every LP (src/pdhg/pdhg.cpp:137-192) with shared parameters and its own norm;
 * p->chunk_attempts is the attempts each LP runs per graph launch (0: 480). */
int pdhg_batch_start(pdhg_batch* h, const pdhg_params* p, const double* spectral_norms);

/* Mixed-precision handoff of the whole batch (no reference counterpart), as pdhg_dev_promote
 * does for one LP: on a PDHG_PRECISION_MIXED batch still in its FP32 phase, widens every LP's
 * iterate pairs to FP64 exactly, keeps every other piece of each LP's state, installs the new
 * tolerance, stopping rule and iteration limit from p, clears the stop flags and recaptures the
 * graph with FP64 values. The next pdhg_batch_run continues every LP in FP64. */
int pdhg_batch_promote(pdhg_batch* h, const pdhg_params* p);

/* Launches `chunks` graph chunks; *running receives how many LPs have not stopped. */
int pdhg_batch_run(pdhg_batch* h, int chunks, int32_t* running);

/* Per-LP engine state (nlp records). */
int pdhg_batch_states(pdhg_batch* h, pdhg_state* out);

/* Report step per LP (as pdhg_dev_finish): col_value and reduced concatenated (sum n), y
 * (sum m), residuals 9 per LP. */
int pdhg_batch_finish(pdhg_batch* h, double* col_value, double* y, double* reduced,
                      double* residuals);

/* The adjoint (no reference counterpart): vector-Jacobian products of each LP's solution map
 * at its reported best point, by implicit differentiation of the PDHG fixed point on its
 * active set with CGLS on the masked matrix, run one workgroup per LP in FP64. The active set
 * is read by complementarity: a column is free when its distance to its nearest bound exceeds
 * |c + A^T y|; a row is active when it is an equality or |y| exceeds its distance to its nearest
 * bound. Inputs, concatenated like the solutions: grad_x = dL/dx (sum n) and grad_y =
 * dL/dy (sum m) for the PDHG multiplier y (the reported row dual is sense * (-y)). Outputs:
 * grad_b (sum m) = dL/d(the active bound of each active row; both bounds of an equality row
 * move together), zero on inactive rows; grad_c (sum n) = dL/d(minimise-space cost), zero on
 * columns at a bound. tol: relative residual of the normal equations at which CGLS stops;
 * iterations and residuals (2 per LP: the x and the y solve; may be NULL). Needs a started
 * batch whose run has ended (pdhg_batch_finish not required). */
int pdhg_batch_adjoint(pdhg_batch* h, const double* grad_x, const double* grad_y, double tol,
                       int max_iterations, double* grad_b, double* grad_c, int32_t* iterations,
                       double* residuals);

/* Test hook: the active set the last pdhg_batch_adjoint call used, 1.0 / 0.0 per coordinate:
 * mask_c (sum n) marks free columns, mask_r (sum m) active rows. */
int pdhg_batch_adjoint_masks(pdhg_batch* h, double* mask_c, double* mask_r);

#ifdef __cplusplus
}
#endif

#endif
