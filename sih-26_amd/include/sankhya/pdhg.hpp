// SPDX-License-Identifier: Apache-2.0
// Restarted PDHG, the first-order LP engine (reference: include/sankhya/pdhg.hpp:9-18).
//
// In this build the engine runs on one MI355X (gfx950) GPU. There is no CPU engine behind
// this entry point: without a usable device the solve returns kNumericalError with a message
// that says so.
#pragma once

#include <vector>

#include "sankhya/logging.hpp"
#include "sankhya/model.hpp"
#include "sankhya/options.hpp"

namespace sankhya::pdhg {

/// Solve an LP with restarted primal-dual hybrid gradient. Never throws.
[[nodiscard]] Solution solve_pdhg(const Model& model, const Options& options, Logger& logger,
                                  SolveControl* control = nullptr);

}  // namespace sankhya::pdhg

namespace sankhya::gpu {

/// The same engine under the name the reference's dispatcher uses for its GPU backend
/// (src/gpu/pdhg_gpu.hpp:19-20).
[[nodiscard]] Solution solve_pdhg_gpu(const Model& model, const Options& options,
                                      Logger& logger, SolveControl* control = nullptr);

/// Many independent LPs solved together by the batched engine (no reference counterpart: the
/// reference solves one LP per call). Each LP runs the same algorithm as solve_pdhg_gpu under
/// the same options; the batch adds a schedule (one workgroup per LP in a persistent launch)
/// and, with pdhg_batch_tune, per-LP workgroup geometries picked by timing. Solutions come
/// back in the order of `models`. Never throws.
[[nodiscard]] std::vector<Solution> solve_pdhg_batch(const std::vector<const Model*>& models,
                                                     const Options& options, Logger& logger,
                                                     SolveControl* control = nullptr);

/// Is a gfx950 device present? Fills `description` (reference src/gpu/device.hpp).
[[nodiscard]] bool device_available(std::string* description);

}  // namespace sankhya::gpu
