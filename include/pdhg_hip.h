/*
 * pdhg_hip.h - the C ABI between the host PDHG driver and the MI355X device engine.
 *
 * The reference has no such layer: its CUDA backend calls cuSPARSE and five small kernels
 * directly from the engine (src/gpu/pdhg_gpu.cu:90-166 kernels, :263-286 SpMV, :297-818
 * driver). Each entry point below replaces one slice of that file, named in its comment.
 *
 * Conventions: plain C types only; every function returns 0 on success or a nonzero code
 * (PDHG_E_*); host pointers are copied and stay owned by the caller; an opaque handle owns
 * all device memory and must be used from the thread that created it. Nothing here falls
 * back to the CPU: without a usable gfx950 device, creation fails with PDHG_E_NO_DEVICE.
 */
#ifndef SANKHYA_PDHG_HIP_H
#define SANKHYA_PDHG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PDHG_OK = 0,
  PDHG_E_NO_DEVICE = 1,   /* no HIP device, or not gfx950 */
  PDHG_E_ALLOC = 2,       /* device or pinned-host allocation failed */
  PDHG_E_LAUNCH = 3,      /* a kernel launch, copy or graph call failed */
  PDHG_E_ARGUMENT = 4,    /* a size or pointer argument is invalid */
  PDHG_E_STATE = 5,       /* call out of order (e.g. run before upload) */
  PDHG_E_INJECTED = 6     /* the fault-injection hook fired (host tests) */
};

/* Precision of the bulk arithmetic. FP64 everywhere, or FP32 matrix values and iterates
 * with FP64 reductions, step control, restart decisions and residuals. */
enum { PDHG_PRECISION_FP64 = 0, PDHG_PRECISION_MIXED = 1 };

typedef struct pdhg_dev pdhg_dev;

/* Replaces device_available() (src/gpu/device.cu:17-41). Writes the device name and
 * architecture into desc. Returns PDHG_OK when device `device` is a gfx950 part. */
int pdhg_dev_probe(int device, char* desc, int desc_len);

/* Replaces the GpuState allocation block (src/gpu/pdhg_gpu.cu:201-258,371-404). */
int pdhg_dev_create(int device, int precision, pdhg_dev** out);
void pdhg_dev_destroy(pdhg_dev* h);

/* Text of the last failure on this handle (never NULL). */
const char* pdhg_dev_last_error(const pdhg_dev* h);

/* Replaces the upload block (src/gpu/pdhg_gpu.cu:406-439) and the cuSPARSE descriptors
 * (:441-493). The scaled matrix Ahat = Dr A Dc is given twice, as CSR (rows of Ahat) and as
 * the CSR of its transpose (columns of Ahat, rows ascending). Vectors, all host pointers:
 *   cost_hat, lower_hat, upper_hat (n): the scaled problem (src/la/scaling.cpp:130-145);
 *   row_lower_hat, row_upper_hat (m);
 *   col_scale (n) = Dc, row_scale (m) = Dr;
 *   min_cost (n): the unscaled minimise-space cost; col_lower/col_upper (n) and
 *   row_lower/row_upper (m): the unscaled bounds, for the convergence test, which is
 *   measured on the original problem (src/pdhg/pdhg_evaluate.cpp:19-139).
 * bound_norm and cost_norm as in src/pdhg/pdhg.cpp:95-111. */
int pdhg_dev_upload(pdhg_dev* h, int32_t m, int32_t n, int64_t nnz, const int32_t* a_rowptr,
                    const int32_t* a_colidx, const double* a_val, const int32_t* at_rowptr,
                    const int32_t* at_colidx, const double* at_val, const double* cost_hat,
                    const double* lower_hat, const double* upper_hat,
                    const double* row_lower_hat, const double* row_upper_hat,
                    const double* col_scale, const double* row_scale, const double* min_cost,
                    const double* col_lower, const double* col_upper, const double* row_lower,
                    const double* row_upper, double bound_norm, double cost_norm);

/* The same upload from the UNSCALED problem, preconditioned on the device: `ruiz_iterations`
 * Ruiz rounds and one Pock-Chambolle round (src/la/scaling.cpp:14-159, build_scaling :111-159),
 * bit-identical to the reference's host passes, applied in place to both copies of the matrix,
 * then the scaled cost and bounds. a_val / at_val hold A's values in CSR / CSC order; min_cost
 * is sense * c; the norms are of the unscaled problem (pdhg.cpp:94-111). extremes (may be
 * null) receives the smallest and largest |entry| of the scaled matrix. */
int pdhg_dev_upload_unscaled(pdhg_dev* h, int32_t m, int32_t n, int64_t nnz,
                             const int32_t* a_rowptr, const int32_t* a_colidx, const double* a_val,
                             const int32_t* at_rowptr, const int32_t* at_colidx,
                             const double* at_val, const double* min_cost, const double* col_lower,
                             const double* col_upper, const double* row_lower,
                             const double* row_upper, double bound_norm, double cost_norm,
                             int ruiz_iterations, int keep_unscaled, double* extremes);

/* With keep_unscaled = 1 the upload above also keeps A's own values in HBM (16 bytes per
 * nonzero more), so that this call can restore them and precondition again, as a new solve of
 * the same LP does, without a new upload from the host. */
int pdhg_dev_precondition(pdhg_dev* h, double* extremes);

/* Test hook: the device's preconditioning results, Dr (m), Dc (n), the scaled values in CSC
 * order (nnz) and the scaled cost and bounds; any pointer may be null. */
int pdhg_dev_scaling(pdhg_dev* h, double* row, double* column, double* csc_values, double* cost_hat,
                     double* col_lower_hat, double* col_upper_hat, double* row_lower_hat,
                     double* row_upper_hat);

/* Replaces estimate_spectral_norm (src/la/scaling.cpp:161-194) on the device: `iterations`
 * power steps on Ahat^T Ahat from the start vector v0 (n, host), rounded up by 1.01. */
int pdhg_dev_spectral_norm(pdhg_dev* h, const double* v0, int iterations, double* norm);

/* Test hooks: y = Ahat x (transpose=0; x has n entries, y m) or y = Ahat^T x (transpose=1),
 * through the same row-binned kernels the iteration uses. Host pointers. Replaces the
 * cuSPARSE calls spmv_nt / spmv_t (src/gpu/pdhg_gpu.cu:263-286). */
int pdhg_dev_spmv(pdhg_dev* h, int transpose, const double* x, double* y);

/* Solver parameters (src/pdhg/pdhg.cpp:119-128,181-186,334). */
typedef struct {
  double tolerance;        /* pdhg_tolerance */
  int64_t iteration_limit; /* already resolved: <0 is not allowed here */
  int32_t restart;         /* pdhg_restart */
  int32_t stop_at_request; /* pdhg_stop_at_request */
  double spectral_norm;    /* eta0 = 1/norm, ceiling = 1e3/max(norm,1e-12) */
  int32_t chunk_attempts;  /* attempt slots per captured graph; 0 = default */
  int32_t stop_rule;       /* 0: the reference's test (request + absolute primal, and the
                              project standard unless stop_at_request); 1: relative KKT
                              error (max of relative primal, dual, gap) <= tolerance */
} pdhg_params;

/* Engine state as seen from the host after a sync. */
typedef struct {
  int64_t iteration;
  int64_t attempts;
  int64_t restarts;
  int64_t ticks;
  int32_t done;       /* 1: stopped on the request (converged) */
  int32_t limit_hit;  /* 1: iteration limit reached */
  int32_t halted;     /* 1: nothing more will run until the host resumes */
  double eta, omega;
  double best_worst;  /* worst() of the best point so far */
  double best_primal, best_dual, best_dual_objective;
} pdhg_state;

/* One evaluation tick (src/pdhg/pdhg.cpp:341-400); the host replays these into its log. */
typedef struct {
  int64_t iteration;
  double primal_objective; /* minimise space, unscaled, no offset */
  double primal, dual, gap;
  int32_t restarted;
  int32_t chose_average;
} pdhg_tick;

/* Iterates at the projection of 0, y = 0, step size from spectral_norm, omega = 1
 * (src/pdhg/pdhg.cpp:137-192). Captures the iteration graph on first use. */
int pdhg_dev_start(pdhg_dev* h, const pdhg_params* p);

/* Replaces the hot loop body (src/gpu/pdhg_gpu.cu:542-744): launches up to `chunks` graph
 * chunks, each a fixed number of attempts plus one evaluation tick, with step control,
 * restarts and the stopping test decided on the device. Returns after they complete, with
 * the state. Ticks produced are appended to an internal log read by pdhg_dev_ticks. */
int pdhg_dev_run(pdhg_dev* h, int chunks, pdhg_state* state);

/* Mixed-precision handoff (no reference counterpart: the reference is FP64 throughout). On a
 * PDHG_PRECISION_MIXED handle still in its FP32 phase: widens the iterate pairs to FP64
 * (exactly), keeps every other piece of state (step size, primal weight, running averages,
 * restart point, best point, counters), installs the new tolerance, stopping rule and
 * iteration limit from p, clears the stop flag, and recaptures the graph chunk with FP64
 * matrix values. The next pdhg_dev_run continues the same run in FP64 (the "FP64 polish").
 * The evaluation ticks and pdhg_dev_finish read FP64 values in both phases. */
int pdhg_dev_promote(pdhg_dev* h, const pdhg_params* p);

/* Copies out tick records with index >= first (at most max); *count receives how many. */
int pdhg_dev_ticks(pdhg_dev* h, int64_t first, pdhg_tick* out, int64_t max, int64_t* count);

/* Report step (src/pdhg/pdhg.cpp:486-501): evaluates the best point on the device and
 * downloads it unscaled. col_value and reduced (n), y (m, the PDHG multiplier before the
 * sign flip). residuals[9] = primal, dual, gap, primal_objective, dual_objective,
 * absolute_primal, absolute_dual, gap_as_verified, complementarity. */
int pdhg_dev_finish(pdhg_dev* h, double* col_value, double* y, double* reduced,
                    double* residuals);

/* Test hook for the convergence test alone: evaluate() of an unscaled point (x n, y m) on
 * the device. activity (m), reduced (n) and residuals[9] as above. Replaces the CPU call of
 * evaluate in src/gpu/pdhg_gpu.cu:658,679. */
int pdhg_dev_evaluate(pdhg_dev* h, const double* x, const double* y, double* activity,
                      double* reduced, double* residuals);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The same evaluate() as
 * pdhg_dev_evaluate, bit for bit, but every argument is a DEVICE pointer (x n, y m in; activity
 * m, reduced n, residuals 9 out), and the partial and segment sums go to a scratch array the
 * library holds per GPU (8 MiB, 1,048,576 doubles, kept until the process ends) instead of the
 * handle's buffers: the call allocates nothing, host memory included. An LP whose work plans
 * need more scratch than that (4 doubles per work item of A and of A^T plus one per long-row
 * segment; some hundred million nonzeros) is refused with PDHG_E_ARGUMENT. Calls on one GPU
 * share its scratch array, so calls that overlap on one GPU race; it returns after the
 * device work has finished. */
int pdhg_dev_evaluate_noalloc(pdhg_dev* h, const double* x, const double* y, double* activity,
                              double* reduced, double* residuals);

/* Timing hook: average device duration in milliseconds of the last `pdhg_dev_run` call's
 * kernels of one kind (0 = primal kernel over Ahat^T, 1 = dual kernel over Ahat), measured
 * with HIP events on the engine's stream around `reps` extra standalone launches. */
int pdhg_dev_time_kernels(pdhg_dev* h, int reps, double* primal_ms, double* dual_ms);

/* Bytes of device memory the handle holds. */
int64_t pdhg_dev_bytes(const pdhg_dev* h);

/* Replaces estimate_pdhg_gpu_memory (src/gpu/gpu_memory.cpp:18-46) before anything is
 * allocated: the device bytes a handle of this precision will hold for this LP, from the two
 * CSR patterns the upload takes (host pointers: a_rowptr m+1 and a_colidx nnz, at_rowptr n+1
 * and at_colidx nnz). It builds the same work plans the upload builds, so the figure is what
 * pdhg_dev_bytes reports after pdhg_dev_upload (unscaled = 0) or pdhg_dev_upload_unscaled
 * (unscaled = 1, keep_unscaled as passed there), plus, for a mixed handle, the FP64 iterate
 * pairs pdhg_dev_promote adds. Host arithmetic only (no device call). Returns -1 on bad
 * arguments. */
int64_t pdhg_dev_estimate_bytes(int32_t m, int32_t n, int64_t nnz, const int32_t* a_rowptr,
                                const int32_t* a_colidx, const int32_t* at_rowptr,
                                const int32_t* at_colidx, int precision, int unscaled,
                                int keep_unscaled);

/* Replaces vram_reserve (src/gpu/gpu_memory.cpp:50-55): the headroom kept free beside the
 * estimate, max(256 MiB, total / 10). */
int64_t pdhg_dev_vram_reserve(int64_t total_bytes);

/* Replaces device_free_memory (the dispatcher's VRAM check, src/core/solve.cpp:751-771):
 * free and total bytes of device `device` (hipMemGetInfo). */
int pdhg_dev_memory(int device, int64_t* free_bytes, int64_t* total_bytes);

/* Fault injection for host tests: make the k-th subsequent device call fail with
 * PDHG_E_INJECTED (k >= 1), or disable with 0. */
void pdhg_dev_inject_fault(pdhg_dev* h, int k);

#ifdef __cplusplus
}
#endif

#endif
