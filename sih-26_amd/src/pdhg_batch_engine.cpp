// SPDX-License-Identifier: Apache-2.0
// Host driver of the batched MI355X PDHG engine: sankhya::gpu::solve_pdhg_batch.
//
// Per LP it does what the single-LP driver does (src/pdhg/pdhg.cpp:76-192 setup, :486-567
// report): validate, minimise-space cost and norms, preconditioning, the CSR of Ahat, the
// power iteration's start vector. The preparation runs on all host cores, one LP per task. The
// device part goes through include/pdhg_batch.h: upload once, per-LP spectral norms,
// optionally the schedule tuner, then graph chunks until every LP has stopped or the time
// limit / interrupt ends the batch.
#include <algorithm>
#include <memory>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/pdhg_batch.h"
#include "engine_common.hpp"
#include "engine_stats.hpp"
#include "sankhya/pdhg.hpp"
#include "sankhya/timer.hpp"

namespace sankhya {
namespace hip_engine {

thread_local std::vector<EngineStats> g_batch_stats;
const std::vector<EngineStats>& last_batch_stats() { return g_batch_stats; }

namespace {

}  // namespace

BatchSession::~BatchSession() { pdhg_batch_destroy(handle); }

std::unique_ptr<BatchSession> open_batch(const std::vector<const Model*>& models,
                                         const Options& options, Logger& logger,
                                         SolveControl* control) {
  auto session = std::make_unique<BatchSession>();
  session->models = models;
  Timer timer;
  const std::size_t nlp = models.size();
  std::vector<Solution>& out = session->solutions;
  out.assign(nlp, Solution{});
  g_batch_stats.assign(nlp, EngineStats{});
  const bool mixed = options.get_string("pdhg_precision") == "mixed";
  const std::string algorithm = mixed ? "pdhg-gpu-batch-mixed" : "pdhg-gpu-batch";
  // Models that fail validation get kModelError and stay out of the device batch.
  std::vector<std::size_t> live;
  for (std::size_t k = 0; k < nlp; ++k) {
    out[k].algorithm = algorithm;
    if (!models[k]) {
      out[k].status = SolveStatus::kModelError;
      out[k].message = "null model";
      continue;
    }
    out[k].allocate_for(*models[k]);
    const std::string text = models[k]->validate();
    if (!text.empty()) {
      out[k].status = SolveStatus::kModelError;
      out[k].message = text;
      continue;
    }
    live.push_back(k);
  }
  if (live.empty()) return session;
  auto fail_all = [&](const std::string& what) {
    for (const std::size_t k : live) {
      out[k].status = SolveStatus::kNumericalError;
      out[k].message = "MI355X batched PDHG: " + what;
      out[k].solve_seconds = timer.elapsed_seconds();
    }
    logger.error("MI355X batched PDHG: {}", what);
    pdhg_batch_destroy(session->handle);
    session->handle = nullptr;
    return std::move(session);
  };
  const int device = engine_device(options, logger);
  char desc[256] = {0};
  if (pdhg_dev_probe(device, desc, sizeof(desc)) != PDHG_OK)
    return fail_all(std::string("no usable gfx950 device (") + desc + ")");

  // ---- Host preparation, one LP per task ----
  const std::size_t nb = live.size();
  std::vector<const Model*> live_models;
  for (const std::size_t k : live) live_models.push_back(models[k]);
  std::vector<Prepared> prep;
  std::vector<std::vector<double>> v0;
  // Norms and the CSR of A on the host, one LP per task; the preconditioner runs on the device
  // for every LP at once (pdhg_batch_upload_unscaled).
  prepare_all(live_models, options.get_int("random_seed"), &prep, &v0, /*scale_on_host=*/false);
  const BatchInput in = make_batch_input(live_models, prep, v0, /*host_scaled=*/false);
  const std::vector<int32_t>& m = in.m;
  const std::vector<int32_t>& n = in.n;
  const double setup_seconds = timer.elapsed_seconds();

  session->live = live;
  pdhg_batch*& dev_h = session->handle;
  int rc = pdhg_batch_create(device, mixed ? PDHG_PRECISION_MIXED : PDHG_PRECISION_FP64, &dev_h);
  if (rc != PDHG_OK) return fail_all("batch handle creation failed (code " + std::to_string(rc) + ")");
  auto dev_error = [&](int code, const char* step) {
    return fail_all(std::string(step) + " failed (code " + std::to_string(code) + "): " +
                    pdhg_batch_last_error(dev_h));
  };
  rc = pdhg_batch_upload_unscaled(dev_h, static_cast<int32_t>(nb), in.m.data(), in.n.data(),
                                  in.nnz.data(), in.a_rp.data(), in.a_ci.data(), in.a_v.data(),
                                  in.t_rp.data(), in.t_ci.data(), in.t_v.data(), in.cost.data(),
                                  in.cl.data(), in.cu.data(), in.rl.data(), in.ru.data(),
                                  in.bound_norm.data(), in.cost_norm.data(), kRuizIterations);
  if (rc != PDHG_OK) return dev_error(rc, "upload");
  std::vector<double> norms(nb, 1.0);
  rc = pdhg_batch_spectral_norms(dev_h, in.v0.data(), kPowerIterations, norms.data());
  if (rc != PDHG_OK) return dev_error(rc, "spectral norms");

  const double tolerance = options.get_double("pdhg_tolerance");
  const std::int64_t limit_option = options.get_int("iteration_limit");
  const Count iteration_limit = limit_option < 0 ? Count{1000000} : static_cast<Count>(limit_option);
  const bool stop_at_request = options.get_bool("pdhg_stop_at_request");
  const double time_limit = options.get_double("time_limit");
  pdhg_params params{};
  params.tolerance = tolerance;
  params.iteration_limit = iteration_limit;
  params.restart = options.get_bool("pdhg_restart") ? 1 : 0;
  params.stop_at_request = stop_at_request ? 1 : 0;
  params.spectral_norm = 1.0;
  params.chunk_attempts = static_cast<int32_t>(options.get_int("pdhg_chunk"));
  params.stop_rule = options.get_string("pdhg_stop_rule") == "relative_kkt" ? 1 : 0;
  logger.info("Solving {} LPs with MI355X batched PDHG on {}, precision {}", nb, desc,
              mixed ? "mixed" : "fp64");
  std::vector<int32_t> geo(nb, -1);
  // The tuner decides from measured times, so a second run may give an LP another schedule,
  // and with it another order of its partial sums. Under `deterministic` every LP stays on
  // the cooperative schedule and the results repeat bit for bit.
  const bool tune = options.get_bool("pdhg_batch_tune") && !options.get_bool("deterministic");
  if (tune) {
    rc = pdhg_batch_tune(dev_h, &params, norms.data(),
                         static_cast<int>(options.get_int("pdhg_batch_tune_attempts")), geo.data());
    if (rc != PDHG_OK) return dev_error(rc, "schedule tuning");
    int count[3] = {0, 0, 0}, coop = 0;
    for (const int32_t g : geo) {
      if (g < 0) ++coop;
      else ++count[std::min(g, 2)];
    }
    logger.info("Schedule: {} cooperative, {} persistent at 256 threads, {} at 512, {} at 1024", coop, count[0], count[1], count[2]);
  }
  // Mixed precision with the FP64 polish (pdhg_mixed_polish), as the single-LP engine runs it:
  // the FP32 phase stops each LP on the relative KKT error at max(pdhg_mixed_switch, tolerance);
  // when every LP has stopped, the whole batch continues in FP64 under the requested rule and
  // tolerance (pdhg_batch_promote). No handoff is needed when the request is a relative KKT
  // error no tighter than the switch.
  const pdhg_params final_params = params;
  bool fp32_phase = mixed && options.get_bool("pdhg_mixed_polish");
  if (fp32_phase) {
    params.tolerance = std::max(tolerance, options.get_double("pdhg_mixed_switch"));
    params.stop_rule = 1;
    if (final_params.stop_rule == 1 && params.tolerance <= tolerance) fp32_phase = false;
  }
  rc = pdhg_batch_start(dev_h, &params, norms.data());
  if (rc != PDHG_OK) return dev_error(rc, "start");
  const double loop_start = timer.elapsed_seconds();
  SolveStatus stop_status = SolveStatus::kIterationLimit;
  int32_t running = static_cast<int32_t>(nb);
  std::vector<std::int64_t> fp32_iterations(nb, 0);
  while (running > 0 || fp32_phase) {
    if (running == 0 && fp32_phase) {
      std::vector<pdhg_state> st32(nb);
      rc = pdhg_batch_states(dev_h, st32.data());
      if (rc != PDHG_OK) return dev_error(rc, "state read");
      for (std::size_t q = 0; q < nb; ++q) fp32_iterations[q] = st32[q].iteration;
      logger.info("batch: FP32 phase ended at {:.3f}s; continuing every LP in FP64",
                  timer.elapsed_seconds() - loop_start);
      rc = pdhg_batch_promote(dev_h, &final_params);
      if (rc != PDHG_OK) return dev_error(rc, "FP64 handoff");
      fp32_phase = false;
      running = static_cast<int32_t>(nb);
      continue;
    }
    if (std::isfinite(time_limit) && timer.elapsed_seconds() >= time_limit) {
      stop_status = SolveStatus::kTimeLimit;
      break;
    }
    if (control && control->interruption_requested()) {
      stop_status = SolveStatus::kInterrupted;
      break;
    }
    const int32_t before = running;
    rc = pdhg_batch_run(dev_h, 1, &running);
    if (rc != PDHG_OK) return dev_error(rc, "iteration");
    if (running != before)
      logger.verbose("batch: {} of {} LPs running at {:.3f}s", running, nb,
                     timer.elapsed_seconds() - loop_start);
  }
  const double loop_seconds = timer.elapsed_seconds() - loop_start;

  // ---- Report, per LP ----
  std::int64_t sum_n = 0, sum_m = 0;
  for (std::size_t q = 0; q < nb; ++q) {
    sum_n += n[q];
    sum_m += m[q];
  }
  std::vector<double> x(static_cast<std::size_t>(sum_n)), y(static_cast<std::size_t>(sum_m)),
      red(static_cast<std::size_t>(sum_n)), res(9 * nb);
  std::vector<pdhg_state> states(nb);
  rc = pdhg_batch_finish(dev_h, x.data(), y.data(), red.data(), res.data());
  if (rc != PDHG_OK) return dev_error(rc, "report");
  rc = pdhg_batch_states(dev_h, states.data());
  if (rc != PDHG_OK) return dev_error(rc, "state read");
  const double total = timer.elapsed_seconds();
  std::int64_t off_n = 0, off_m = 0;
  for (std::size_t q = 0; q < nb; ++q) {
    const std::size_t k = live[q];
    Solution& s = out[k];
    std::copy(x.begin() + off_n, x.begin() + off_n + n[q], s.col_value.begin());
    std::vector<double> yl(y.begin() + off_m, y.begin() + off_m + m[q]);
    std::vector<double> rl2(red.begin() + off_n, red.begin() + off_n + n[q]);
    const pdhg_state& st = states[q];
    s.solve_seconds = total;
    EngineStats& es = g_batch_stats[k];
    es.iterations = st.iteration;
    es.fp32_iterations = mixed ? (fp32_iterations[q] > 0 ? fp32_iterations[q] : st.iteration) : 0;
    es.attempts = st.attempts;
    es.restarts = st.restarts;
    es.ticks = st.ticks;
    es.device_bytes = pdhg_batch_bytes(dev_h);
    es.spectral_norm = norms[q];
    es.setup_seconds = setup_seconds;
    es.loop_seconds = loop_seconds;
    es.total_seconds = total;
    es.final_primal = res[9 * q + 0];
    es.final_dual = res[9 * q + 1];
    es.final_gap = res[9 * q + 2];
    es.final_absolute_primal = res[9 * q + 5];
    es.schedule = geo[q];
    report(*models[k], res.data() + 9 * q, yl, rl2,
           // An LP that had already stopped on its own iteration limit keeps that status when
           // a time limit or an interrupt later ends the rest of the batch.
           ReportState{st.done != 0,
                       st.done        ? SolveStatus::kOptimal
                       : st.limit_hit ? SolveStatus::kIterationLimit
                                      : stop_status,
                       st.iteration, st.restarts, tolerance, stop_at_request, final_params.stop_rule == 1},
           &s);
    off_n += n[q];
    off_m += m[q];
  }
  logger.info("Batch of {} LPs: setup {:.3f}s, iterations {:.3f}s, total {:.3f}s", nb, setup_seconds,
              loop_seconds, total);
  return session;
}

std::vector<Solution> run_batch(const std::vector<const Model*>& models, const Options& options,
                                Logger& logger, SolveControl* control) {
  return open_batch(models, options, logger, control)->solutions;
}

int BatchSession::adjoint_masks(double* mask_c, double* mask_r) const {
  std::vector<std::int64_t> off_n(models.size() + 1, 0), off_m(models.size() + 1, 0);
  for (std::size_t k = 0; k < models.size(); ++k) {
    off_n[k + 1] = off_n[k] + (models[k] ? models[k]->num_cols() : 0);
    off_m[k + 1] = off_m[k] + (models[k] ? models[k]->num_rows() : 0);
  }
  std::fill(mask_c, mask_c + off_n.back(), 0.0);
  std::fill(mask_r, mask_r + off_m.back(), 0.0);
  if (!handle || live.empty()) return PDHG_E_STATE;
  std::size_t sn = 0, sm = 0;
  for (const std::size_t k : live) {
    sn += static_cast<std::size_t>(off_n[k + 1] - off_n[k]);
    sm += static_cast<std::size_t>(off_m[k + 1] - off_m[k]);
  }
  std::vector<double> mc(sn), mr(sm);
  const int rc = pdhg_batch_adjoint_masks(handle, mc.data(), mr.data());
  if (rc != PDHG_OK) return rc;
  std::size_t pn = 0, pm = 0;
  for (const std::size_t k : live) {
    const std::int64_t n = off_n[k + 1] - off_n[k], m = off_m[k + 1] - off_m[k];
    std::copy(mc.begin() + static_cast<std::ptrdiff_t>(pn), mc.begin() + static_cast<std::ptrdiff_t>(pn + n),
              mask_c + off_n[k]);
    std::copy(mr.begin() + static_cast<std::ptrdiff_t>(pm), mr.begin() + static_cast<std::ptrdiff_t>(pm + m),
              mask_r + off_m[k]);
    pn += static_cast<std::size_t>(n);
    pm += static_cast<std::size_t>(m);
  }
  return PDHG_OK;
}

int BatchSession::adjoint(const double* grad_x, const double* grad_row_dual, double tol,
                          int max_iterations, double* grad_c, double* grad_bound,
                          int32_t* iterations, double* residuals) const {
  // Model order -> device order (LPs that failed validation are not on the device).
  std::vector<std::int64_t> off_n(models.size() + 1, 0), off_m(models.size() + 1, 0);
  for (std::size_t k = 0; k < models.size(); ++k) {
    off_n[k + 1] = off_n[k] + (models[k] ? models[k]->num_cols() : 0);
    off_m[k + 1] = off_m[k] + (models[k] ? models[k]->num_rows() : 0);
  }
  std::fill(grad_c, grad_c + off_n.back(), 0.0);
  std::fill(grad_bound, grad_bound + off_m.back(), 0.0);
  if (iterations) std::fill(iterations, iterations + 2 * models.size(), 0);
  if (residuals) std::fill(residuals, residuals + 2 * models.size(), 0.0);
  if (!handle || live.empty()) return PDHG_E_STATE;
  std::vector<double> gx, gy;
  for (const std::size_t k : live) {
    const double sense = models[k]->sense_multiplier();
    gx.insert(gx.end(), grad_x + off_n[k], grad_x + off_n[k + 1]);
    // row_dual = sense * (-y)  =>  dL/dy = -sense * dL/d(row_dual)
    for (std::int64_t i = off_m[k]; i < off_m[k + 1]; ++i) gy.push_back(-sense * grad_row_dual[i]);
  }
  std::vector<double> gb(gy.size()), gc(gx.size()), res(2 * live.size());
  std::vector<int32_t> its(2 * live.size());
  const int rc = pdhg_batch_adjoint(handle, gx.data(), gy.data(), tol, max_iterations, gb.data(),
                                    gc.data(), its.data(), res.data());
  if (rc != PDHG_OK) return rc;
  std::size_t pn = 0, pm = 0;
  for (std::size_t q = 0; q < live.size(); ++q) {
    const std::size_t k = live[q];
    const double sense = models[k]->sense_multiplier();
    const std::int64_t n = off_n[k + 1] - off_n[k], m = off_m[k + 1] - off_m[k];
    // The engine's cost is sense * c.
    for (std::int64_t j = 0; j < n; ++j) grad_c[off_n[k] + j] = sense * gc[pn + j];
    for (std::int64_t i = 0; i < m; ++i) grad_bound[off_m[k] + i] = gb[pm + i];
    if (iterations) {
      iterations[2 * k] = its[2 * q];
      iterations[2 * k + 1] = its[2 * q + 1];
    }
    if (residuals) {
      residuals[2 * k] = res[2 * q];
      residuals[2 * k + 1] = res[2 * q + 1];
    }
    pn += static_cast<std::size_t>(n);
    pm += static_cast<std::size_t>(m);
  }
  return PDHG_OK;
}

}  // namespace hip_engine

namespace gpu {
std::vector<Solution> solve_pdhg_batch(const std::vector<const Model*>& models,
                                       const Options& options, Logger& logger,
                                       SolveControl* control) {
  try {
    return hip_engine::run_batch(models, options, logger, control);
  } catch (const std::exception& e) {
    std::vector<Solution> out(models.size());
    for (Solution& s : out) {
      s.status = SolveStatus::kNumericalError;
      s.algorithm = "pdhg-gpu-batch";
      s.message = std::string("batched PDHG failed: ") + e.what();
    }
    return out;
  }
}
}  // namespace gpu

}  // namespace sankhya
