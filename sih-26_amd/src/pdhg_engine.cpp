// SPDX-License-Identifier: Apache-2.0
// Host driver of the MI355X restarted-PDHG engine: sankhya::pdhg::solve_pdhg and
// sankhya::gpu::solve_pdhg_gpu.
//
// Reference behaviour followed: src/pdhg/pdhg.cpp:76-568 (problem setup, preconditioning, the
// iteration's decisions, the report) and src/gpu/pdhg_gpu.cu:297-818 (the same on a GPU). What
// differs is where the work happens: the iteration, the convergence test, restarts and the
// stopping test all run on the device (include/pdhg_hip.h), and the host only launches
// captured graph chunks, replays the evaluation ticks into the log, and enforces the time
// limit, interrupts and the progress callback between chunks. A device failure ends the solve
// with kNumericalError; there is no CPU engine behind it.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "engine_common.hpp"
#include "engine_stats.hpp"
#include "sankhya/pdhg.hpp"
#include "sankhya/timer.hpp"

namespace sankhya {
namespace hip_engine {

thread_local EngineStats g_last_stats;

const EngineStats& last_engine_stats() { return g_last_stats; }

namespace {

struct DevHandle {
  pdhg_dev* h = nullptr;
  ~DevHandle() { pdhg_dev_destroy(h); }
};

}  // namespace


Solution run(const Model& model, const Options& options, Logger& logger, SolveControl* control) {
  Timer timer;
  g_last_stats = EngineStats{};
  Solution solution;
  solution.allocate_for(model);
  const bool mixed = options.get_string("pdhg_precision") == "mixed";
  solution.algorithm = mixed ? "pdhg-gpu-mixed" : "pdhg-gpu";

  const std::string problem_text = model.validate();
  if (!problem_text.empty()) {
    solution.status = SolveStatus::kModelError;
    solution.message = problem_text;
    return solution;
  }
  auto device_failure = [&](const std::string& what) {
    solution.status = SolveStatus::kNumericalError;
    solution.message = "MI355X PDHG: " + what;
    solution.solve_seconds = timer.elapsed_seconds();
    logger.error("{}", solution.message);
    return solution;
  };

  const int device = engine_device(options, logger);
  char desc[256] = {0};
  if (pdhg_dev_probe(device, desc, sizeof(desc)) != PDHG_OK)
    return device_failure(std::string("no usable gfx950 device (") + desc + ")");

  const Index rows = model.num_rows(), cols = model.num_cols();
  const double sense = model.sense_multiplier();
  const auto n = static_cast<std::size_t>(cols), m = static_cast<std::size_t>(rows);

  const double t_checked = timer.elapsed_seconds();
  // Host side: minimise-space cost, norms and the CSR of A; the preconditioner runs on the
  // device (pdhg_dev_upload_unscaled).
  const Prepared prep = prepare_unscaled(model, host_threads());
  const std::vector<double>& cost = prep.cost;
  g_last_stats.setup_seconds = timer.elapsed_seconds();

  // The VRAM check before anything is allocated (src/core/solve.cpp:751-771,
  // src/gpu/gpu_memory.cpp:18-55). The reference then falls back to its CPU engine; this
  // engine has none, so an LP that does not fit ends here with the figures in the message.
  const int precision = mixed ? PDHG_PRECISION_MIXED : PDHG_PRECISION_FP64;
  {
    const std::int64_t required = pdhg_dev_estimate_bytes(
        rows, cols, model.num_nonzeros(), prep.a_rp.data(), prep.a_ci.data(),
        model.matrix.column_starts().data(), model.matrix.row_indices().data(), precision,
        /*unscaled=*/1, /*keep_unscaled=*/0);
    std::int64_t free_bytes = 0, total_bytes = 0;
    if (pdhg_dev_memory(device, &free_bytes, &total_bytes) == PDHG_OK) {
      // Test hook: pretend the device has this many bytes free.
      if (const char* f = std::getenv("SANKHYA_HIP_FREE_BYTES")) free_bytes = std::atoll(f);
      const MemoryCheck mc = check_device_memory(required, free_bytes, total_bytes);
      logger.info("GPU PDHG memory: required {:.0f} MiB, available {:.0f} MiB, reserve {:.0f} MiB",
                  mc.required_mib, mc.free_mib, mc.reserve_mib);
      if (!mc.fits) return device_failure(mc.message);
    }
  }

  DevHandle dev;
  int rc = pdhg_dev_create(device, precision, &dev.h);
  if (rc != PDHG_OK) return device_failure("device handle creation failed (code " + std::to_string(rc) + ")");
  const int fault = static_cast<int>(std::getenv("SANKHYA_HIP_FAIL_AT") ? std::atoi(std::getenv("SANKHYA_HIP_FAIL_AT")) : 0);
  if (fault > 0) pdhg_dev_inject_fault(dev.h, fault);
  auto dev_error = [&](int code, const char* step) {
    return device_failure(std::string(step) + " failed (code " + std::to_string(code) + "): " +
                          pdhg_dev_last_error(dev.h));
  };
  double extremes[2] = {0.0, 0.0};
  const double t_scale0 = timer.elapsed_seconds();
  rc = pdhg_dev_upload_unscaled(dev.h, rows, cols, model.num_nonzeros(), prep.a_rp.data(),
                                prep.a_ci.data(), prep.a_v.data(), model.matrix.column_starts().data(),
                                model.matrix.row_indices().data(), model.matrix.values().data(),
                                cost.data(), model.col_lower.data(), model.col_upper.data(),
                                model.row_lower.data(), model.row_upper.data(), prep.bound_norm,
                                prep.cost_norm, kRuizIterations, /*keep_unscaled=*/0, extremes);
  if (rc != PDHG_OK) return dev_error(rc, "upload");

  const std::vector<double> v0 = power_start(model, options.get_int("random_seed"));
  double spectral_norm = 1.0;
  const double t_uploaded = timer.elapsed_seconds();
  rc = pdhg_dev_spectral_norm(dev.h, v0.data(), kPowerIterations, &spectral_norm);
  if (rc != PDHG_OK) return dev_error(rc, "spectral norm");
  const double t_norm = timer.elapsed_seconds();
  g_last_stats.spectral_norm = spectral_norm;

  const double tolerance = options.get_double("pdhg_tolerance");
  const std::int64_t limit_option = options.get_int("iteration_limit");
  const Count iteration_limit = limit_option < 0 ? Count{1000000} : static_cast<Count>(limit_option);
  const bool use_restarts = options.get_bool("pdhg_restart");
  const bool stop_at_request = options.get_bool("pdhg_stop_at_request");
  const double time_limit = options.get_double("time_limit");
  const bool has_time_limit = std::isfinite(time_limit);

  logger.info("Solving LP with MI355X restarted PDHG on {}: {} rows, {} columns, {} nonzeros", desc,
              rows, cols, model.num_nonzeros());
  logger.info("Scaled matrix entries in [{:.3e}, {:.3e}], estimated ||A||_2 = {:.4e}", extremes[0],
              extremes[1], spectral_norm);
  logger.info("Target relative tolerance {:.1e}, restarts {}, precision {}", tolerance,
              use_restarts ? "on" : "off", mixed ? "mixed" : "fp64");

  pdhg_params params{};
  params.tolerance = tolerance;
  params.iteration_limit = iteration_limit;
  params.restart = use_restarts ? 1 : 0;
  params.stop_at_request = stop_at_request ? 1 : 0;
  params.spectral_norm = spectral_norm;
  params.chunk_attempts = static_cast<int32_t>(options.get_int("pdhg_chunk"));
  params.stop_rule = options.get_string("pdhg_stop_rule") == "relative_kkt" ? 1 : 0;
  // Mixed precision with the FP64 polish: the FP32 phase stops on the relative KKT error at
  // max(pdhg_mixed_switch, tolerance); the same run then continues in FP64 under the requested
  // rule and tolerance (pdhg_dev_promote). A stalled FP32 phase hands over early.
  const pdhg_params final_params = params;
  bool fp32_phase = mixed && options.get_bool("pdhg_mixed_polish");
  if (fp32_phase) {
    params.tolerance = std::max(tolerance, options.get_double("pdhg_mixed_switch"));
    params.stop_rule = 1;
  }
  rc = pdhg_dev_start(dev.h, &params);
  if (rc != PDHG_OK) return dev_error(rc, "start");
  const double loop_start = timer.elapsed_seconds();
  logger.debug("Setup: validate and probe {:.4f}s, CSR {:.4f}s, device handle {:.4f}s, upload and "
               "device preconditioning {:.4f}s, spectral norm {:.4f}s, start and graph capture "
               "{:.4f}s", t_checked, prep.csr_seconds, t_scale0 - g_last_stats.setup_seconds,
               t_uploaded - t_scale0, t_norm - t_uploaded, loop_start - t_norm);

  pdhg_state st{};
  SolveStatus stop_status = SolveStatus::kIterationLimit;
  bool converged = false, logged_table = false;
  std::int64_t ticks_seen = 0;
  std::vector<pdhg_tick> tick_buf(4096);
  int batch = 1;
  double stall_best = std::numeric_limits<double>::infinity();
  std::int64_t stall_tick = 0;
  // FP32 phase stall: no 10% gain in the best KKT error for max(100, ticks so far / 2) ticks
  // (progress comes in bursts at restarts, whose spacing grows with the iteration count).
  constexpr std::int64_t kStallTicks = 100;
  auto promote = [&]() {
    g_last_stats.fp32_iterations = st.iteration;
    logger.info("FP32 phase ended at iteration {} (relative KKT {:.3e}); continuing in FP64",
                st.iteration, st.best_worst);
    fp32_phase = false;
    batch = 1;
    return pdhg_dev_promote(dev.h, &final_params);
  };
  while (true) {
    const double elapsed = timer.elapsed_seconds();
    if (has_time_limit && elapsed >= time_limit) {
      stop_status = SolveStatus::kTimeLimit;
      break;
    }
    if (control) {
      if (control->interruption_requested()) {
        stop_status = SolveStatus::kInterrupted;
        break;
      }
      if (control->progress_callback && control->callback_due()) {
        Progress p;
        p.phase = Progress::Phase::kLp;
        p.iterations = st.iteration;
        p.objective = st.best_primal;
        p.best_bound = sense * st.best_dual_objective + model.objective_offset;
        p.elapsed_seconds = elapsed;
        if (control->progress_callback(p) != 0 || control->interruption_requested()) {
          control->interrupt();
          stop_status = SolveStatus::kInterrupted;
          break;
        }
      }
    }
    const double t0 = timer.elapsed_seconds();
    rc = pdhg_dev_run(dev.h, batch, &st);
    if (rc != PDHG_OK) return dev_error(rc, "iteration");
    const double dt = timer.elapsed_seconds() - t0;
    // Replay the evaluation ticks into the iteration table (pdhg.cpp:386-400).
    while (ticks_seen < st.ticks) {
      std::int64_t got = 0;
      rc = pdhg_dev_ticks(dev.h, ticks_seen, tick_buf.data(), static_cast<std::int64_t>(tick_buf.size()), &got);
      if (rc != PDHG_OK) return dev_error(rc, "tick read");
      if (got == 0) break;
      for (std::int64_t k = 0; k < got; ++k) {
        if (!logged_table) {
          logger.begin_iteration_table();
          logged_table = true;
        }
        const pdhg_tick& t = tick_buf[k];
        logger.iteration(t.iteration, sense * t.primal_objective + model.objective_offset, t.primal,
                         t.dual, timer.elapsed_seconds());
      }
      ticks_seen += got;
    }
    if (fp32_phase && st.done) {
      // Already there when the request is a relative KKT error no tighter than the switch.
      if (final_params.stop_rule == 1 && params.tolerance <= tolerance) {
        g_last_stats.fp32_iterations = st.iteration;
        converged = true;
        break;
      }
      rc = promote();
      if (rc != PDHG_OK) return dev_error(rc, "FP64 handoff");
      continue;
    }
    if (fp32_phase && !st.limit_hit) {
      if (st.best_worst < 0.9 * stall_best) {
        stall_best = st.best_worst;
        stall_tick = st.ticks;
      } else if (st.ticks - stall_tick >= std::max(kStallTicks, st.ticks / 2)) {
        rc = promote();
        if (rc != PDHG_OK) return dev_error(rc, "FP64 handoff");
        continue;
      }
    }
    if (st.done) {
      converged = true;
      break;
    }
    if (st.limit_hit || st.halted) break;
    // Keep each host wait near 20 ms so limits and interrupts stay responsive.
    if (dt < 0.01 && batch < 256) batch *= 2;
    else if (dt > 0.04 && batch > 1) batch /= 2;
  }
  g_last_stats.loop_seconds = timer.elapsed_seconds() - loop_start;

  // ---- Report (pdhg.cpp:486-567) ----
  std::vector<double> y(m), reduced(n);
  double res[9];
  rc = pdhg_dev_finish(dev.h, solution.col_value.data(), y.data(), reduced.data(), res);
  if (rc != PDHG_OK) return dev_error(rc, "report");
  const Residuals final_r = from_array(res);
  g_last_stats.iterations = st.iteration;
  g_last_stats.attempts = st.attempts;
  g_last_stats.restarts = st.restarts;
  g_last_stats.ticks = st.ticks;
  g_last_stats.device_bytes = pdhg_dev_bytes(dev.h);
  g_last_stats.final_primal = final_r.primal;
  g_last_stats.final_dual = final_r.dual;
  g_last_stats.final_gap = final_r.gap;
  g_last_stats.final_absolute_primal = final_r.absolute_primal;
  solution.solve_seconds = timer.elapsed_seconds();
  report(model, res, y, reduced,
         ReportState{converged, stop_status, st.iteration, st.restarts, tolerance, stop_at_request,
                     final_params.stop_rule == 1},
         &solution, &prep);
  const double t_report = timer.elapsed_seconds();
  pdhg_dev_destroy(dev.h);
  dev.h = nullptr;
  g_last_stats.total_seconds = timer.elapsed_seconds();
  logger.debug("Teardown: finish and report {:.4f}s, device release {:.4f}s",
               t_report - loop_start - g_last_stats.loop_seconds, g_last_stats.total_seconds - t_report);

  logger.info("");
  logger.info("Status: {}   objective {:.10e}   iterations {}   restarts {}   time {:.3f}s",
              to_string(solution.status), solution.objective, solution.iterations, st.restarts,
              solution.solve_seconds);
  logger.info("Relative residuals: primal {:.3e}, dual {:.3e}, gap {:.3e}", final_r.primal,
              final_r.dual, final_r.gap);
  if (!solution.message.empty()) logger.info("{}", solution.message);
  return solution;
}

}  // namespace hip_engine

namespace pdhg {
Solution solve_pdhg(const Model& model, const Options& options, Logger& logger, SolveControl* control) {
  return hip_engine::run(model, options, logger, control);
}
}  // namespace pdhg

namespace gpu {
Solution solve_pdhg_gpu(const Model& model, const Options& options, Logger& logger, SolveControl* control) {
  return hip_engine::run(model, options, logger, control);
}

bool device_available(std::string* description) {
  char buf[256] = {0};
  const int rc = pdhg_dev_probe(0, buf, sizeof(buf));
  if (description) *description = buf;
  return rc == PDHG_OK;
}
}  // namespace gpu

}  // namespace sankhya
