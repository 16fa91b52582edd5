/*
 * sankhya_c.h - SANKHYA's own C API (include/sankhya/sankhya.h) for the one path this engine
 * replaces: build an LP, set options by name, solve, read the solution. Each function keeps the
 * reference's name, arguments, return codes and contract (src/api/c_api.cpp); a caller written
 * against SANKHYA's C API links this library instead and runs the MI355X PDHG engine.
 *
 * What is left out: CPLEX LP and QPS reading, quadratic objectives, warm starts from a basis
 * (sankhya_solve_from), basis statuses, certificates, ranging, IIS and the solution pool, which
 * belong to SANKHYA's other engines. The model handle is the same sankhya_model as in
 * sankhya_pdhg.h, so a model made from arrays (sankhya_model_new) can be solved here too.
 *
 * Every engine option of the registry is accepted by the sankhya_options_set_* calls, with the
 * reference's names; `gpu` is accepted and PDHG runs on the MI355X either way (this library has
 * no CPU engine).
 */
#ifndef SANKHYA_C_H
#define SANKHYA_C_H

#include <stdint.h>

#include "sankhya_pdhg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/sankhya/sankhya.h:42-51 */
typedef enum sankhya_status {
  SANKHYA_OK = 0,
  SANKHYA_ERROR_ARGUMENT = 1,
  SANKHYA_ERROR_IO = 2,
  SANKHYA_ERROR_MODEL = 3,
  SANKHYA_ERROR_OPTION = 4,
  SANKHYA_ERROR_MEMORY = 5,
  SANKHYA_ERROR_INTERNAL = 6
} sankhya_status;

/* include/sankhya/sankhya.h:53-69 (values are the reference's) */
typedef enum sankhya_solve_status {
  SANKHYA_NOT_SOLVED = 0,
  SANKHYA_OPTIMAL = 1,
  SANKHYA_FEASIBLE = 2,
  SANKHYA_INFEASIBLE = 3,
  SANKHYA_UNBOUNDED = 4,
  SANKHYA_INFEASIBLE_OR_UNBOUNDED = 10,
  SANKHYA_ITERATION_LIMIT = 5,
  SANKHYA_TIME_LIMIT = 6,
  SANKHYA_NODE_LIMIT = 7,
  SANKHYA_NUMERICAL_ERROR = 8,
  SANKHYA_MODEL_ERROR = 9,
  SANKHYA_INTERRUPTED = 11
} sankhya_solve_status;

typedef struct sankhya_options sankhya_options;
typedef struct sankhya_solution sankhya_solution;

/* sankhya.h:81-91 */
const char* sankhya_version(void);
const char* sankhya_last_error(void); /* reason for the last failing call on this thread */
double sankhya_infinity(void);

/* sankhya.h:96-188: an empty model, then columns, rows and coefficients. */
sankhya_model* sankhya_model_create(void);
/* sankhya.h:99-106: read an MPS file (free or fixed, plain or gzip-compressed) into the
 * handle, replacing it. A `.lp` file or a quadratic section fails with SANKHYA_ERROR_IO; on any failure the
 * handle keeps what it held and sankhya_last_error() names the file and line. */
sankhya_status sankhya_model_read(sankhya_model* model, const char* path);
sankhya_status sankhya_model_set_maximize(sankhya_model* model, int maximize);
sankhya_status sankhya_model_set_objective_offset(sankhya_model* model, double offset);
sankhya_status sankhya_model_add_column(sankhya_model* model, double cost, double lower,
                                        double upper, int is_integer, const char* name,
                                        int* index);
sankhya_status sankhya_model_add_row(sankhya_model* model, double lower, double upper,
                                     const char* name, int* index);
/* Setting the same entry twice replaces it; an exact zero removes it (sankhya.h:133-141). */
sankhya_status sankhya_model_set_coefficient(sankhya_model* model, int row, int col,
                                             double value);
sankhya_status sankhya_model_set_col_bounds(sankhya_model* model, int col, double lower,
                                            double upper);
sankhya_status sankhya_model_set_row_bounds(sankhya_model* model, int row, double lower,
                                            double upper);
sankhya_status sankhya_model_set_objective_coefficient(sankhya_model* model, int col,
                                                       double cost);
int sankhya_model_num_cols(const sankhya_model* model);
int sankhya_model_num_rows(const sankhya_model* model);
int sankhya_model_num_nonzeros(const sankhya_model* model);
sankhya_status sankhya_model_validate(const sankhya_model* model);

/* sankhya.h:192-220: progress and interruption (the SolveControl of the solve). */
typedef enum sankhya_progress_phase {
  SANKHYA_PHASE_PRESOLVE = 0,
  SANKHYA_PHASE_LP = 1,
  SANKHYA_PHASE_TREE = 2
} sankhya_progress_phase;

typedef struct sankhya_progress {
  sankhya_progress_phase phase;
  int64_t iterations;
  int64_t nodes;
  double objective;  /* PDHG: the best point's relative primal residual (the reference's quirk) */
  double best_bound; /* PDHG: sense * best dual objective + offset */
  double gap;
  double elapsed_seconds;
  int64_t open_nodes;
} sankhya_progress;

/* A callback returning non-zero interrupts the solve at the next chunk boundary. */
sankhya_status sankhya_set_callback(sankhya_model* model,
                                    int (*callback)(const sankhya_progress*, void*),
                                    void* user_data);
/* From any thread, during a solve of `model`. */
sankhya_status sankhya_model_interrupt(sankhya_model* model);

/* sankhya.h:225-233 */
sankhya_options* sankhya_options_create(void);
void sankhya_options_free(sankhya_options* options);
sankhya_status sankhya_options_set_bool(sankhya_options* options, const char* name, int value);
sankhya_status sankhya_options_set_int(sankhya_options* options, const char* name,
                                       int64_t value);
sankhya_status sankhya_options_set_double(sankhya_options* options, const char* name,
                                          double value);
sankhya_status sankhya_options_set_string(sankhya_options* options, const char* name,
                                          const char* value);

/* sankhya.h:248: the return value says whether the CALL succeeded; the solution's status says
 * what the solver concluded. `options` may be NULL for the defaults. */
sankhya_status sankhya_solve(sankhya_model* model, const sankhya_options* options,
                             sankhya_solution** solution);

/* sankhya.h:268-358 */
void sankhya_solution_free(sankhya_solution* solution);
sankhya_solve_status sankhya_solution_status(const sankhya_solution* solution);
const char* sankhya_solution_message(const sankhya_solution* solution);
const char* sankhya_solution_algorithm(const sankhya_solution* solution);
double sankhya_solution_objective(const sankhya_solution* solution);
double sankhya_solution_dual_bound(const sankhya_solution* solution);
double sankhya_solution_relative_gap(const sankhya_solution* solution);
int64_t sankhya_solution_iterations(const sankhya_solution* solution);
double sankhya_solution_seconds(const sankhya_solution* solution);
double sankhya_solution_primal_infeasibility(const sankhya_solution* solution);
double sankhya_solution_dual_infeasibility(const sankhya_solution* solution);
/* `count` must equal the model's column (row) count, else SANKHYA_ERROR_ARGUMENT. */
sankhya_status sankhya_solution_col_values(const sankhya_solution* solution, double* values,
                                           int count);
sankhya_status sankhya_solution_row_activities(const sankhya_solution* solution, double* values,
                                               int count);
sankhya_status sankhya_solution_row_duals(const sankhya_solution* solution, double* values,
                                          int count);
sankhya_status sankhya_solution_col_duals(const sankhya_solution* solution, double* values,
                                          int count);
int sankhya_solution_claims_a_point(const sankhya_solution* solution);

#ifdef __cplusplus
}
#endif

#endif
