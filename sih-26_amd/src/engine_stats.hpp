// SPDX-License-Identifier: Apache-2.0
// Counters of the last engine run on this thread, for the C ABI and the benchmark.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/pdhg_batch.h"
#include "sankhya/logging.hpp"
#include "sankhya/model.hpp"
#include "sankhya/options.hpp"

namespace sankhya::hip_engine {

struct EngineStats {
  std::int64_t iterations = 0, attempts = 0, restarts = 0, ticks = 0, device_bytes = 0;
  std::int64_t fp32_iterations = 0;  // mixed precision: iterations before the FP64 handoff
  std::int32_t schedule = -1;  // batched engine: -1 cooperative, else the persistent geometry
  double spectral_norm = 0.0;
  double setup_seconds = 0.0, loop_seconds = 0.0, total_seconds = 0.0;
  double final_primal = 0.0, final_dual = 0.0, final_gap = 0.0;
  double final_absolute_primal = 0.0;
};

const EngineStats& last_engine_stats();
// Per-LP counters of the last batched run on this thread, in model order.
const std::vector<EngineStats>& last_batch_stats();

Solution run(const Model& model, const Options& options, Logger& logger, SolveControl* control);
std::vector<Solution> run_batch(const std::vector<const Model*>& models, const Options& options,
                                Logger& logger, SolveControl* control);

// A solved batch that keeps its device state for vector-Jacobian products.
struct BatchSession {
  std::vector<const Model*> models;
  std::vector<std::size_t> live;  // models on the device, in device order
  std::vector<Solution> solutions;
  ::pdhg_batch* handle = nullptr;
  ~BatchSession();
  // dL/dc and dL/d(active row bound) for upstream dL/dx and dL/d(row_dual), concatenated in
  // model order (user-facing signs: the model's own sense and row duals).
  int adjoint(const double* grad_x, const double* grad_row_dual, double tol, int max_iterations,
              double* grad_c, double* grad_bound, int32_t* iterations, double* residuals) const;
  // The active set of the last adjoint, in model order (1 free column / active row, else 0).
  int adjoint_masks(double* mask_c, double* mask_r) const;
};
std::unique_ptr<BatchSession> open_batch(const std::vector<const Model*>& models,
                                         const Options& options, Logger& logger,
                                         SolveControl* control);

}  // namespace sankhya::hip_engine
