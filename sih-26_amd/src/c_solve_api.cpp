// SPDX-License-Identifier: Apache-2.0
// SANKHYA's C API (include/sankhya_c.h) over this engine's solve(). Reference counterpart:
// src/api/c_api.cpp, whose contract each function keeps: every failure returns a code and
// leaves its reason in a thread-local last error, nothing throws across the boundary, and the
// solve runs with a SolveControl that sankhya_model_interrupt can reach from another thread.
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sankhya_c.h"
#include "c_model.hpp"
#include "sankhya/io.hpp"
#include "sankhya/options.hpp"
#include "sankhya/pdhg.hpp"

struct sankhya_options {
  sankhya::Options options;
};

struct sankhya_solution {
  sankhya::Solution solution;
};

using namespace sankhya;

namespace {

thread_local std::string g_last_error;

sankhya_status fail(sankhya_status code, const std::string& why) {
  g_last_error = why;
  return code;
}

sankhya_status ok() { return SANKHYA_OK; }

// Runs `body`, turning any exception into a status (src/api/c_api.cpp:131-160).
template <typename Body>
sankhya_status guarded(Body&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(SANKHYA_ERROR_MEMORY, "out of memory");
  } catch (const std::exception& e) {
    return fail(SANKHYA_ERROR_INTERNAL, e.what());
  } catch (...) {
    return fail(SANKHYA_ERROR_INTERNAL, "unknown exception");
  }
}

sankhya_solve_status to_c_status(SolveStatus status) {
  switch (status) {
    case SolveStatus::kNotSolved: return SANKHYA_NOT_SOLVED;
    case SolveStatus::kOptimal: return SANKHYA_OPTIMAL;
    case SolveStatus::kFeasible: return SANKHYA_FEASIBLE;
    case SolveStatus::kInfeasible: return SANKHYA_INFEASIBLE;
    case SolveStatus::kUnbounded: return SANKHYA_UNBOUNDED;
    case SolveStatus::kInfeasibleOrUnbounded: return SANKHYA_INFEASIBLE_OR_UNBOUNDED;
    case SolveStatus::kIterationLimit: return SANKHYA_ITERATION_LIMIT;
    case SolveStatus::kTimeLimit: return SANKHYA_TIME_LIMIT;
    case SolveStatus::kNodeLimit: return SANKHYA_NODE_LIMIT;
    case SolveStatus::kNumericalError: return SANKHYA_NUMERICAL_ERROR;
    case SolveStatus::kModelError: return SANKHYA_MODEL_ERROR;
    case SolveStatus::kInterrupted: return SANKHYA_INTERRUPTED;
  }
  return SANKHYA_NOT_SOLVED;
}

// The builder calls edit the pending entries; a model made from arrays moves its matrix there
// on its first edit.
void make_pending(sankhya_model* h) {
  if (h->pending) return;
  const SparseMatrix& a = h->model.matrix;
  h->entries.clear();
  for (Index j = 0; j < a.num_cols(); ++j)
    for (Index k = a.column_starts()[j]; k < a.column_starts()[j + 1]; ++k)
      h->entries[{a.row_indices()[k], j}] = a.values()[k];
  h->pending = true;
}

// The model as a solve sees it: the pending entries materialised (rows ascending per column).
Model built_model(const sankhya_model* h) {
  Model md = h->model;
  if (!h->pending) return md;
  md.matrix.reset(md.num_rows(), md.num_cols());
  md.matrix.reserve(h->entries.size());
  for (const auto& [rc, v] : h->entries) md.matrix.add_entry(rc.first, rc.second, v);
  md.matrix.finalize(0.0);
  return md;
}

bool row_ok(const sankhya_model* h, int row) { return row >= 0 && row < h->model.num_rows(); }
bool col_ok(const sankhya_model* h, int col) { return col >= 0 && col < h->model.num_cols(); }

sankhya_status set_option(sankhya_options* options, const char* name, const std::string& text) {
  if (!options || !name) return fail(SANKHYA_ERROR_ARGUMENT, "null options or name");
  std::string error;
  if (!options->options.set_from_string(name, text, &error)) return fail(SANKHYA_ERROR_OPTION, error);
  return ok();
}

sankhya_status copy_vector(const std::vector<double>& source, double* destination, int count,
                           const char* what) {
  if (!destination) return fail(SANKHYA_ERROR_ARGUMENT, "destination is null");
  if (count < 0 || static_cast<std::size_t>(count) != source.size())
    return fail(SANKHYA_ERROR_ARGUMENT, std::string("wrong buffer size for ") + what +
                                            ": the solution has " + std::to_string(source.size()) +
                                            " entries, the buffer " + std::to_string(count));
  if (count > 0) std::memcpy(destination, source.data(), sizeof(double) * source.size());
  return ok();
}

}  // namespace

extern "C" {

const char* sankhya_version(void) { return "sankhya-mi355x PDHG engine (gfx950)"; }
const char* sankhya_last_error(void) { return g_last_error.c_str(); }
double sankhya_infinity(void) { return std::numeric_limits<double>::infinity(); }

sankhya_model* sankhya_model_create(void) {
  try {
    auto* h = new sankhya_model();
    h->model.matrix.reset(0, 0);
    h->model.matrix.finalize(0.0);
    return h;
  } catch (...) {
    g_last_error = "out of memory";
    return nullptr;
  }
}

// src/api/c_api.cpp:185-200: on failure the handle keeps what it held.
sankhya_status sankhya_model_read(sankhya_model* h, const char* path) {
  if (!h || !path) return fail(SANKHYA_ERROR_ARGUMENT, "model or path is null");
  return guarded([&] {
    Model fresh;
    const io::ReadResult result = io::read_model(path, &fresh);
    if (!result.ok) return fail(SANKHYA_ERROR_IO, result.error);
    h->model = std::move(fresh);
    h->entries.clear();
    h->pending = false;
    return ok();
  });
}

sankhya_status sankhya_model_set_maximize(sankhya_model* h, int maximize) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  h->model.sense = maximize ? ObjSense::kMaximize : ObjSense::kMinimize;
  return ok();
}

sankhya_status sankhya_model_set_objective_offset(sankhya_model* h, double offset) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  h->model.objective_offset = offset;
  return ok();
}

sankhya_status sankhya_model_add_column(sankhya_model* h, double cost, double lower, double upper,
                                        int is_integer, const char* name, int* index) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  return guarded([&] {
    make_pending(h);
    Model& md = h->model;
    const Index j = md.num_cols();
    md.col_cost.push_back(cost);
    md.col_lower.push_back(lower);
    md.col_upper.push_back(upper);
    md.col_type.push_back(is_integer ? VarType::kInteger : VarType::kContinuous);
    if (name || !md.col_names.empty()) {
      md.col_names.resize(static_cast<std::size_t>(j));
      md.col_names.push_back(name ? name : "C" + std::to_string(j));
    }
    if (index) *index = j;
    return ok();
  });
}

sankhya_status sankhya_model_add_row(sankhya_model* h, double lower, double upper,
                                     const char* name, int* index) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  return guarded([&] {
    make_pending(h);
    Model& md = h->model;
    const Index i = md.num_rows();
    md.row_lower.push_back(lower);
    md.row_upper.push_back(upper);
    if (name || !md.row_names.empty()) {
      md.row_names.resize(static_cast<std::size_t>(i));
      md.row_names.push_back(name ? name : "R" + std::to_string(i));
    }
    if (index) *index = i;
    return ok();
  });
}

sankhya_status sankhya_model_set_coefficient(sankhya_model* h, int row, int col, double value) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  if (!row_ok(h, row) || !col_ok(h, col))
    return fail(SANKHYA_ERROR_ARGUMENT, "coefficient (" + std::to_string(row) + ", " +
                                            std::to_string(col) + ") is outside the model");
  return guarded([&] {
    make_pending(h);
    if (value == 0.0) h->entries.erase({row, col});
    else h->entries[{row, col}] = value;
    return ok();
  });
}

sankhya_status sankhya_model_set_col_bounds(sankhya_model* h, int col, double lower, double upper) {
  if (!h || !col_ok(h, col)) return fail(SANKHYA_ERROR_ARGUMENT, "no such column");
  h->model.col_lower[static_cast<std::size_t>(col)] = lower;
  h->model.col_upper[static_cast<std::size_t>(col)] = upper;
  return ok();
}

sankhya_status sankhya_model_set_row_bounds(sankhya_model* h, int row, double lower, double upper) {
  if (!h || !row_ok(h, row)) return fail(SANKHYA_ERROR_ARGUMENT, "no such row");
  h->model.row_lower[static_cast<std::size_t>(row)] = lower;
  h->model.row_upper[static_cast<std::size_t>(row)] = upper;
  return ok();
}

sankhya_status sankhya_model_set_objective_coefficient(sankhya_model* h, int col, double cost) {
  if (!h || !col_ok(h, col)) return fail(SANKHYA_ERROR_ARGUMENT, "no such column");
  h->model.col_cost[static_cast<std::size_t>(col)] = cost;
  return ok();
}

int sankhya_model_num_cols(const sankhya_model* h) { return h ? h->model.num_cols() : 0; }
int sankhya_model_num_rows(const sankhya_model* h) { return h ? h->model.num_rows() : 0; }
int sankhya_model_num_nonzeros(const sankhya_model* h) {
  if (!h) return 0;
  return h->pending ? static_cast<int>(h->entries.size()) : h->model.num_nonzeros();
}

sankhya_status sankhya_model_validate(const sankhya_model* h) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  return guarded([&] {
    const std::string problem = built_model(h).validate();
    if (!problem.empty()) return fail(SANKHYA_ERROR_MODEL, problem);
    return ok();
  });
}

sankhya_status sankhya_set_callback(sankhya_model* h, int (*callback)(const sankhya_progress*, void*),
                                    void* user_data) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  if (!callback) {
    h->progress_callback = nullptr;
    return ok();
  }
  h->progress_callback = [callback, user_data](const Progress& p) {
    sankhya_progress c{};
    c.phase = p.phase == Progress::Phase::kPresolve ? SANKHYA_PHASE_PRESOLVE : SANKHYA_PHASE_LP;
    c.iterations = p.iterations;
    c.objective = p.objective;
    c.best_bound = p.best_bound;
    c.elapsed_seconds = p.elapsed_seconds;
    return callback(&c, user_data);
  };
  return ok();
}

sankhya_status sankhya_model_interrupt(sankhya_model* h) {
  if (!h) return fail(SANKHYA_ERROR_ARGUMENT, "null model");
  std::lock_guard<std::mutex> lock(h->control_mutex);
  if (h->active_control) h->active_control->interrupt();
  return ok();
}

sankhya_options* sankhya_options_create(void) {
  try {
    return new sankhya_options();
  } catch (...) {
    g_last_error = "out of memory";
    return nullptr;
  }
}

void sankhya_options_free(sankhya_options* options) { delete options; }

sankhya_status sankhya_options_set_bool(sankhya_options* o, const char* name, int value) {
  return set_option(o, name, value ? "true" : "false");
}
sankhya_status sankhya_options_set_int(sankhya_options* o, const char* name, int64_t value) {
  return set_option(o, name, std::to_string(value));
}
sankhya_status sankhya_options_set_double(sankhya_options* o, const char* name, double value) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", value);
  return set_option(o, name, buf);
}
sankhya_status sankhya_options_set_string(sankhya_options* o, const char* name, const char* value) {
  if (!value) return fail(SANKHYA_ERROR_ARGUMENT, "null value");
  return set_option(o, name, value);
}

sankhya_status sankhya_solve(sankhya_model* h, const sankhya_options* options,
                             sankhya_solution** solution) {
  if (!h || !solution) return fail(SANKHYA_ERROR_ARGUMENT, "null model or solution pointer");
  *solution = nullptr;
  return guarded([&] {
    const Model built = built_model(h);
    const Options effective = options ? options->options : Options();
    auto control = std::make_shared<SolveControl>();
    control->progress_callback = h->progress_callback;
    {
      std::lock_guard<std::mutex> lock(h->control_mutex);
      h->active_control = control;
    }
    struct Clear {
      sankhya_model* m;
      ~Clear() {
        std::lock_guard<std::mutex> lock(m->control_mutex);
        m->active_control.reset();
      }
    } clear{h};
    auto* result = new sankhya_solution();
    result->solution = sankhya::solve(built, effective, control.get());
    *solution = result;
    return ok();
  });
}

void sankhya_solution_free(sankhya_solution* s) { delete s; }

sankhya_solve_status sankhya_solution_status(const sankhya_solution* s) {
  return s ? to_c_status(s->solution.status) : SANKHYA_NOT_SOLVED;
}
const char* sankhya_solution_message(const sankhya_solution* s) { return s ? s->solution.message.c_str() : ""; }
const char* sankhya_solution_algorithm(const sankhya_solution* s) { return s ? s->solution.algorithm.c_str() : ""; }
double sankhya_solution_objective(const sankhya_solution* s) { return s ? s->solution.objective : 0.0; }
double sankhya_solution_dual_bound(const sankhya_solution* s) { return s ? s->solution.dual_bound : 0.0; }
double sankhya_solution_relative_gap(const sankhya_solution* s) { return s ? s->solution.relative_gap : 0.0; }
int64_t sankhya_solution_iterations(const sankhya_solution* s) {
  return s ? static_cast<int64_t>(s->solution.iterations) : 0;
}
double sankhya_solution_seconds(const sankhya_solution* s) { return s ? s->solution.solve_seconds : 0.0; }
double sankhya_solution_primal_infeasibility(const sankhya_solution* s) {
  return s ? s->solution.primal_infeasibility : 0.0;
}
double sankhya_solution_dual_infeasibility(const sankhya_solution* s) {
  return s ? s->solution.dual_infeasibility : 0.0;
}
sankhya_status sankhya_solution_col_values(const sankhya_solution* s, double* values, int count) {
  if (!s) return fail(SANKHYA_ERROR_ARGUMENT, "null solution");
  return copy_vector(s->solution.col_value, values, count, "column values");
}
sankhya_status sankhya_solution_row_activities(const sankhya_solution* s, double* values, int count) {
  if (!s) return fail(SANKHYA_ERROR_ARGUMENT, "null solution");
  return copy_vector(s->solution.row_activity, values, count, "row activities");
}
sankhya_status sankhya_solution_row_duals(const sankhya_solution* s, double* values, int count) {
  if (!s) return fail(SANKHYA_ERROR_ARGUMENT, "null solution");
  return copy_vector(s->solution.row_dual, values, count, "row duals");
}
sankhya_status sankhya_solution_col_duals(const sankhya_solution* s, double* values, int count) {
  if (!s) return fail(SANKHYA_ERROR_ARGUMENT, "null solution");
  return copy_vector(s->solution.col_dual, values, count, "column duals");
}
int sankhya_solution_claims_a_point(const sankhya_solution* s) {
  return s && claims_a_point(s->solution.status) ? 1 : 0;
}

}  // extern "C"
