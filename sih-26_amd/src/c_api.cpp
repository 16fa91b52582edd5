// SPDX-License-Identifier: Apache-2.0
// C ABI for foreign callers (include/sankhya_pdhg.h). Reference counterpart: the C API over
// solve(), src/api/c_api.cpp:563, which likewise catches everything and returns codes.
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <vector>
#include <string>

#include "../../include/sankhya_pdhg.h"
#include "c_model.hpp"
#include "engine_common.hpp"
#include "engine_stats.hpp"
#include "generate.hpp"
#include "sankhya/io.hpp"
#include "sankhya/pdhg.hpp"
#include "scaling.hpp"


using namespace sankhya;

namespace {

void copy_text(char* dst, std::size_t cap, const std::string& s) {
  const std::size_t k = std::min(cap - 1, s.size());
  std::memcpy(dst, s.data(), k);
  dst[k] = '\0';
}

bool apply_options(Options* o, const char* text, std::string* error) {
  if (!text) return true;
  std::string s(text);
  std::size_t pos = 0;
  while (pos < s.size()) {
    std::size_t end = s.find_first_of(";,", pos);
    if (end == std::string::npos) end = s.size();
    const std::string item = s.substr(pos, end - pos);
    pos = end + 1;
    if (item.find_first_not_of(" \t") == std::string::npos) continue;
    const std::size_t eq = item.find('=');
    if (eq == std::string::npos) {
      *error = "option without '=': " + item;
      return false;
    }
    auto trim = [](std::string t) {
      const auto b = t.find_first_not_of(" \t");
      const auto e = t.find_last_not_of(" \t");
      return b == std::string::npos ? std::string() : t.substr(b, e - b + 1);
    };
    if (!o->set_from_string(trim(item.substr(0, eq)), trim(item.substr(eq + 1)), error)) return false;
  }
  return true;
}

}  // namespace

extern "C" {

sankhya_model* sankhya_model_new(int32_t m, int32_t n, int64_t nnz, const int32_t* colptr,
                                 const int32_t* rowidx, const double* val, const double* cost,
                                 const double* col_lower, const double* col_upper,
                                 const double* row_lower, const double* row_upper,
                                 double objective_offset, int32_t maximize) {
  if (m < 0 || n < 0 || nnz < 0 || !colptr || colptr[0] != 0 || colptr[n] != nnz) return nullptr;
  for (int32_t j = 0; j < n; ++j) {
    if (colptr[j + 1] < colptr[j]) return nullptr;
    for (int32_t k = colptr[j]; k < colptr[j + 1]; ++k) {
      if (rowidx[k] < 0 || rowidx[k] >= m) return nullptr;
      if (k > colptr[j] && rowidx[k] <= rowidx[k - 1]) return nullptr;
    }
  }
  auto* h = new sankhya_model();
  Model& md = h->model;
  md.resize_columns(n);
  md.resize_rows(m);
  for (int32_t j = 0; j < n; ++j) {
    md.col_cost[j] = cost ? cost[j] : 0.0;
    md.col_lower[j] = col_lower ? normalize_infinity(col_lower[j]) : 0.0;
    md.col_upper[j] = col_upper ? normalize_infinity(col_upper[j]) : kInfinity;
  }
  for (int32_t i = 0; i < m; ++i) {
    md.row_lower[i] = row_lower ? normalize_infinity(row_lower[i]) : -kInfinity;
    md.row_upper[i] = row_upper ? normalize_infinity(row_upper[i]) : kInfinity;
  }
  md.objective_offset = objective_offset;
  md.sense = maximize ? ObjSense::kMaximize : ObjSense::kMinimize;
  md.matrix.assign_csc(m, n, std::vector<Index>(colptr, colptr + n + 1),
                       std::vector<Index>(rowidx, rowidx + nnz), std::vector<double>(val, val + nnz));
  return h;
}

sankhya_model* sankhya_model_generate(int32_t kind, int32_t m, int32_t n, int32_t per_col,
                                      double param, uint64_t seed) {
  if (kind < 0 || kind > 2 || m <= 0 || n <= 0 || per_col < 0) return nullptr;
  hip_engine::GenSpec g;
  g.kind = kind;
  g.m = m;
  g.n = n;
  g.per_col = per_col;
  g.param = param;
  g.seed = seed;
  try {
    auto* h = new sankhya_model();
    h->model = hip_engine::generate(g);
    return h;
  } catch (...) {
    return nullptr;
  }
}

void sankhya_model_dims(const sankhya_model* h, int32_t* m, int32_t* n, int64_t* nnz) {
  if (m) *m = h->model.num_rows();
  if (n) *n = h->model.num_cols();
  if (nnz) *nnz = h->model.num_nonzeros();
}

void sankhya_model_export(const sankhya_model* h, int32_t* colptr, int32_t* rowidx, double* val,
                          double* cost, double* col_lower, double* col_upper, double* row_lower,
                          double* row_upper, double* objective_offset, int32_t* maximize) {
  const Model& md = h->model;
  auto cp = [](auto* dst, const auto& src) {
    if (dst) std::memcpy(dst, src.data(), sizeof(src[0]) * src.size());
  };
  cp(colptr, md.matrix.column_starts());
  cp(rowidx, md.matrix.row_indices());
  cp(val, md.matrix.values());
  cp(cost, md.col_cost);
  cp(col_lower, md.col_lower);
  cp(col_upper, md.col_upper);
  cp(row_lower, md.row_lower);
  cp(row_upper, md.row_upper);
  if (objective_offset) *objective_offset = md.objective_offset;
  if (maximize) *maximize = md.sense == ObjSense::kMaximize ? 1 : 0;
}

void sankhya_model_free(sankhya_model* h) { delete h; }

int sankhya_pdhg_solve(const sankhya_model* h, const char* options, double* col_value,
                       double* row_dual, double* col_dual, double* row_activity,
                       sankhya_result* out) {
  Options o;
  o.set_bool("log_to_console", false);
  std::string error;
  if (!apply_options(&o, options, &error)) {
    if (out) {
      std::memset(out, 0, sizeof(*out));
      out->status = static_cast<int32_t>(SolveStatus::kModelError);
      copy_text(out->message, sizeof(out->message), error);
    }
    return -1;
  }
  const Solution s = solve(h->model, o, nullptr);
  const auto& st = hip_engine::last_engine_stats();
  auto cp = [](double* dst, const std::vector<double>& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(double) * src.size());
  };
  cp(col_value, s.col_value);
  cp(row_dual, s.row_dual);
  cp(col_dual, s.col_dual);
  cp(row_activity, s.row_activity);
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->status = static_cast<int32_t>(s.status);
    out->iterations = s.iterations;
    out->attempts = st.attempts;
    out->restarts = st.restarts;
    out->ticks = st.ticks;
    out->device_bytes = st.device_bytes;
    out->fp32_iterations = st.fp32_iterations;
    out->schedule = -1;
    out->objective = s.objective;
    out->dual_bound = s.dual_bound;
    out->solve_seconds = s.solve_seconds;
    out->setup_seconds = st.setup_seconds;
    out->loop_seconds = st.loop_seconds;
    out->spectral_norm = st.spectral_norm;
    out->primal_infeasibility = s.primal_infeasibility;
    out->dual_infeasibility = s.dual_infeasibility;
    out->complementarity_violation = s.complementarity_violation;
    out->relative_gap = s.relative_gap;
    out->final_primal = st.final_primal;
    out->final_dual = st.final_dual;
    out->final_gap = st.final_gap;
    out->final_absolute_primal = st.final_absolute_primal;
    copy_text(out->algorithm, sizeof(out->algorithm), s.algorithm);
    copy_text(out->message, sizeof(out->message), s.message);
  }
  return 0;
}

struct sankhya_batch {
  std::unique_ptr<hip_engine::BatchSession> session;
};

sankhya_batch* sankhya_batch_solve(const sankhya_model* const* models, int32_t nlp,
                                   const char* options, double* col_value, double* row_dual,
                                   double* col_dual, double* row_activity,
                                   sankhya_result* results) {
  Options o;
  o.set_bool("log_to_console", false);
  std::string error;
  if (!apply_options(&o, options, &error) || nlp < 0 || (nlp > 0 && !models)) {
    if (results && nlp > 0) {
      std::memset(results, 0, sizeof(*results));
      results->status = static_cast<int32_t>(SolveStatus::kModelError);
      copy_text(results->message, sizeof(results->message), error.empty() ? "bad batch" : error);
    }
    return nullptr;
  }
  std::vector<const Model*> list(static_cast<std::size_t>(nlp));
  for (int32_t k = 0; k < nlp; ++k) list[k] = models[k] ? &models[k]->model : nullptr;
  Logger logger(o.get_bool("log_to_console") ? stdout : nullptr, log_level_of(o.get_string("log_level")));
  auto* out_session = new sankhya_batch();
  try {
    out_session->session = hip_engine::open_batch(list, o, logger, nullptr);
  } catch (const std::exception& e) {
    out_session->session = std::make_unique<hip_engine::BatchSession>();
    out_session->session->models = list;
    out_session->session->solutions.assign(list.size(), Solution{});
    for (Solution& s : out_session->session->solutions) {
      s.status = SolveStatus::kNumericalError;
      s.message = std::string("batched PDHG failed: ") + e.what();
    }
  }
  const std::vector<Solution>& sols = out_session->session->solutions;
  const auto& stats = hip_engine::last_batch_stats();
  std::size_t on = 0, om = 0;
  for (int32_t k = 0; k < nlp; ++k) {
    const Solution& s = sols[k];
    const std::size_t n = models[k] ? static_cast<std::size_t>(models[k]->model.num_cols()) : 0;
    const std::size_t m = models[k] ? static_cast<std::size_t>(models[k]->model.num_rows()) : 0;
    auto cp = [](double* dst, std::size_t off, const std::vector<double>& src, std::size_t len) {
      if (dst && src.size() == len && len) std::memcpy(dst + off, src.data(), sizeof(double) * len);
    };
    cp(col_value, on, s.col_value, n);
    cp(col_dual, on, s.col_dual, n);
    cp(row_dual, om, s.row_dual, m);
    cp(row_activity, om, s.row_activity, m);
    on += n;
    om += m;
    if (!results) continue;
    sankhya_result* out = results + k;
    std::memset(out, 0, sizeof(*out));
    out->status = static_cast<int32_t>(s.status);
    out->iterations = s.iterations;
    if (static_cast<std::size_t>(k) < stats.size()) {
      const auto& st = stats[k];
      out->attempts = st.attempts;
      out->restarts = st.restarts;
      out->ticks = st.ticks;
      out->device_bytes = st.device_bytes;
      out->setup_seconds = st.setup_seconds;
      out->loop_seconds = st.loop_seconds;
      out->spectral_norm = st.spectral_norm;
      out->final_primal = st.final_primal;
      out->final_dual = st.final_dual;
      out->final_gap = st.final_gap;
      out->final_absolute_primal = st.final_absolute_primal;
      out->schedule = st.schedule;
      out->fp32_iterations = st.fp32_iterations;
    }
    out->objective = s.objective;
    out->dual_bound = s.dual_bound;
    out->solve_seconds = s.solve_seconds;
    out->primal_infeasibility = s.primal_infeasibility;
    out->dual_infeasibility = s.dual_infeasibility;
    out->complementarity_violation = s.complementarity_violation;
    out->relative_gap = s.relative_gap;
    copy_text(out->algorithm, sizeof(out->algorithm), s.algorithm);
    copy_text(out->message, sizeof(out->message), s.message);
  }
  return out_session;
}

int sankhya_batch_adjoint(const sankhya_batch* session, const double* grad_col_value,
                          const double* grad_row_dual, double tol, int max_iterations,
                          double* grad_cost, double* grad_row_bound, int32_t* iterations,
                          double* residuals) {
  if (!session || !session->session || !grad_col_value || !grad_row_dual || !grad_cost ||
      !grad_row_bound)
    return PDHG_E_ARGUMENT;
  return session->session->adjoint(grad_col_value, grad_row_dual, tol, max_iterations, grad_cost,
                                   grad_row_bound, iterations, residuals);
}

int sankhya_batch_adjoint_masks(const sankhya_batch* session, double* mask_c, double* mask_r) {
  if (!session || !session->session || !mask_c || !mask_r) return PDHG_E_ARGUMENT;
  return session->session->adjoint_masks(mask_c, mask_r);
}

void sankhya_batch_free(sankhya_batch* session) { delete session; }

int sankhya_pdhg_solve_batch(const sankhya_model* const* models, int32_t nlp, const char* options,
                             double* col_value, double* row_dual, double* col_dual,
                             double* row_activity, sankhya_result* results) {
  sankhya_batch* s = sankhya_batch_solve(models, nlp, options, col_value, row_dual, col_dual,
                                         row_activity, results);
  if (!s) return -1;
  sankhya_batch_free(s);
  return 0;
}

int sankhya_model_set_names(sankhya_model* h, const char* name, const char* source_path,
                            const char* const* col_names, const char* const* row_names) {
  if (!h) return -1;
  if (name) h->model.name = name;
  if (source_path) h->model.source_path = source_path;
  auto fill = [](std::vector<std::string>* out, const char* const* names, Index count) {
    if (!names) return;
    out->assign(static_cast<std::size_t>(count), std::string());
    for (Index k = 0; k < count; ++k)
      if (names[k]) (*out)[static_cast<std::size_t>(k)] = names[k];
  };
  fill(&h->model.col_names, col_names, h->model.num_cols());
  fill(&h->model.row_names, row_names, h->model.num_rows());
  return 0;
}

const char* sankhya_model_name_of(const sankhya_model* h, int32_t kind, int32_t index) {
  if (!h) return nullptr;
  const Model& md = h->model;
  auto pick = [index](const std::vector<std::string>& names) -> const char* {
    if (index < 0 || static_cast<std::size_t>(index) >= names.size()) return nullptr;
    return names[static_cast<std::size_t>(index)].c_str();
  };
  switch (kind) {
    case 0: return md.name.c_str();
    case 1: return md.source_path.c_str();
    case 2: return pick(md.col_names);
    case 3: return pick(md.row_names);
    default: return nullptr;
  }
}

int32_t sankhya_model_col_is_integer(const sankhya_model* h, int32_t col) {
  if (!h || col < 0 || static_cast<std::size_t>(col) >= h->model.col_type.size()) return -1;
  return h->model.col_type[static_cast<std::size_t>(col)] == VarType::kInteger ? 1 : 0;
}

int sankhya_write_solution(const sankhya_model* h, const char* options, const char* path,
                           const sankhya_result* result, const double* col_value,
                           const double* row_dual, const double* col_dual,
                           const double* row_activity, char* error, int32_t error_len) {
  std::string err;
  auto fail = [&](const std::string& e) {
    if (error && error_len > 0) copy_text(error, static_cast<std::size_t>(error_len), e);
    return -1;
  };
  if (!h || !path || !result) return fail("model, path and result are required");
  Options o;
  if (!apply_options(&o, options, &err)) return fail(err);
  try {
    const Index n = h->model.num_cols(), m = h->model.num_rows();
    Solution s;
    s.status = static_cast<SolveStatus>(result->status);
    s.objective = result->objective;
    s.dual_bound = result->dual_bound;
    s.iterations = result->iterations;
    s.solve_seconds = result->solve_seconds;
    s.primal_infeasibility = result->primal_infeasibility;
    s.dual_infeasibility = result->dual_infeasibility;
    s.algorithm.assign(result->algorithm, strnlen(result->algorithm, sizeof(result->algorithm)));
    s.message.assign(result->message, strnlen(result->message, sizeof(result->message)));
    auto take = [](std::vector<double>* dst, const double* src, Index count) {
      if (src) dst->assign(src, src + count);
    };
    take(&s.col_value, col_value, n);
    take(&s.col_dual, col_dual, n);
    take(&s.row_dual, row_dual, m);
    take(&s.row_activity, row_activity, m);
    if (!io::write_solution(path, h->model, s, o, &err)) return fail(err);
  } catch (const std::exception& e) {
    return fail(e.what());
  }
  return 0;
}

int sankhya_pdhg_scaling(const sankhya_model* h, int ruiz, int threads, double* row, double* column,
                         double* scaled_val, double* cost_hat, double* col_lower_hat,
                         double* col_upper_hat, double* row_lower_hat, double* row_upper_hat) {
  const Model& md = h->model;
  std::vector<double> cost(md.col_cost.size());
  for (std::size_t j = 0; j < cost.size(); ++j) cost[j] = md.sense_multiplier() * md.col_cost[j];
  const hip_engine::Scaling s = hip_engine::build_scaling(md, cost, ruiz, threads);
  auto cp = [](double* dst, const std::vector<double>& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(double) * src.size());
  };
  cp(row, s.row);
  cp(column, s.column);
  cp(scaled_val, s.values);
  cp(cost_hat, s.cost);
  cp(col_lower_hat, s.col_lower);
  cp(col_upper_hat, s.col_upper);
  cp(row_lower_hat, s.row_lower);
  cp(row_upper_hat, s.row_upper);
  return 0;
}

int sankhya_pdhg_power_start(const sankhya_model* h, int64_t random_seed, double* v0) {
  if (!h || (!v0 && h->model.num_cols() > 0)) return PDHG_E_ARGUMENT;
  const std::vector<double> v = hip_engine::power_start(h->model, random_seed);
  if (!v.empty()) std::memcpy(v0, v.data(), sizeof(double) * v.size());
  return 0;
}

int sankhya_pdhg_csr(const sankhya_model* h, int threads, int32_t* rowptr, int32_t* colidx,
                     double* values) {
  if (!h || !rowptr) return PDHG_E_ARGUMENT;
  const hip_engine::Prepared p = hip_engine::prepare(h->model, threads);
  std::memcpy(rowptr, p.a_rp.data(), sizeof(int32_t) * p.a_rp.size());
  if (colidx && !p.a_ci.empty()) std::memcpy(colidx, p.a_ci.data(), sizeof(int32_t) * p.a_ci.size());
  if (values && !p.a_v.empty()) std::memcpy(values, p.a_v.data(), sizeof(double) * p.a_v.size());
  return 0;
}

int sankhya_pdhg_device_open(const sankhya_model* h, int device, int precision, pdhg_dev** out) {
  if (!h || !out) return PDHG_E_ARGUMENT;
  const Model& md = h->model;
  // As the engine does it: cost, norms and the CSR of A on the host, the preconditioner on
  // the device.
  const hip_engine::Prepared p = hip_engine::prepare_unscaled(md, hip_engine::host_threads());
  int rc = pdhg_dev_create(device, precision, out);
  if (rc) return rc;
  rc = pdhg_dev_upload_unscaled(*out, md.num_rows(), md.num_cols(), md.num_nonzeros(),
                                p.a_rp.data(), p.a_ci.data(), p.a_v.data(),
                                md.matrix.column_starts().data(), md.matrix.row_indices().data(),
                                md.matrix.values().data(), p.cost.data(), md.col_lower.data(),
                                md.col_upper.data(), md.row_lower.data(), md.row_upper.data(),
                                p.bound_norm, p.cost_norm, hip_engine::kRuizIterations,
                                /*keep_unscaled=*/1, nullptr);
  if (rc) {
    pdhg_dev_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

int sankhya_pdhg_batch_open(const sankhya_model* const* models, int32_t nlp, int device,
                            int precision, pdhg_batch** out) {
  if (!models || nlp <= 0 || !out) return PDHG_E_ARGUMENT;
  *out = nullptr;
  std::vector<const Model*> ms;
  for (int32_t k = 0; k < nlp; ++k) {
    if (!models[k]) return PDHG_E_ARGUMENT;
    ms.push_back(&models[k]->model);
  }
  std::vector<hip_engine::Prepared> prep;
  std::vector<std::vector<double>> v0;
  hip_engine::prepare_all(ms, 0, &prep, &v0, /*scale_on_host=*/false);
  const hip_engine::BatchInput in = hip_engine::make_batch_input(ms, prep, v0, /*host_scaled=*/false);
  int rc = pdhg_batch_create(device, precision, out);
  if (rc) return rc;
  rc = pdhg_batch_upload_unscaled(*out, nlp, in.m.data(), in.n.data(), in.nnz.data(), in.a_rp.data(),
                                  in.a_ci.data(), in.a_v.data(), in.t_rp.data(), in.t_ci.data(),
                                  in.t_v.data(), in.cost.data(), in.cl.data(), in.cu.data(),
                                  in.rl.data(), in.ru.data(), in.bound_norm.data(),
                                  in.cost_norm.data(), hip_engine::kRuizIterations);
  if (rc) {
    pdhg_batch_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

int sankhya_pdhg_quality(const sankhya_model* h, const double* col_value, const double* row_dual,
                         const double* col_dual, double dual_bound, double* row_activity,
                         double* out) {
  if (!h || !out) return -1;
  const Model& md = h->model;
  const auto n = static_cast<std::size_t>(md.num_cols()), m = static_cast<std::size_t>(md.num_rows());
  if ((n && (!col_value || !col_dual)) || (m && (!row_dual || !row_activity))) return -1;
  Solution s;
  s.col_value.assign(col_value, col_value + n);
  s.row_dual.assign(row_dual, row_dual + m);
  s.col_dual.assign(col_dual, col_dual + n);
  s.dual_bound = dual_bound;
  s.recompute_quality(md);
  if (m) std::memcpy(row_activity, s.row_activity.data(), sizeof(double) * m);
  const double vals[9] = {s.objective,          s.primal_infeasibility,
                          s.primal_infeasibility_scaled, s.dual_infeasibility,
                          s.dual_infeasibility_scaled,   s.complementarity_violation,
                          s.integrality_violation,       s.absolute_gap,
                          s.relative_gap};
  std::memcpy(out, vals, sizeof(vals));
  return 0;
}

int sankhya_pdhg_quality_threaded(const sankhya_model* h, const double* col_value,
                                  const double* row_dual, const double* col_dual, double dual_bound,
                                  int threads, double* row_activity, double* out) {
  if (!h || !out || threads < 1) return -1;
  const Model& md = h->model;
  const auto n = static_cast<std::size_t>(md.num_cols()), m = static_cast<std::size_t>(md.num_rows());
  if ((n && (!col_value || !col_dual)) || (m && (!row_dual || !row_activity))) return -1;
  try {
    const hip_engine::Prepared prep = hip_engine::prepare_unscaled(md, 1);
    Solution s;
    s.col_value.assign(col_value, col_value + n);
    s.row_dual.assign(row_dual, row_dual + m);
    s.col_dual.assign(col_dual, col_dual + n);
    s.dual_bound = dual_bound;
    hip_engine::recompute_quality_threaded(md, prep.a_rp, prep.a_ci, prep.a_v, threads, &s);
    if (m) std::memcpy(row_activity, s.row_activity.data(), sizeof(double) * m);
    const double vals[9] = {s.objective,          s.primal_infeasibility,
                            s.primal_infeasibility_scaled, s.dual_infeasibility,
                            s.dual_infeasibility_scaled,   s.complementarity_violation,
                            s.integrality_violation,       s.absolute_gap,
                            s.relative_gap};
    std::memcpy(out, vals, sizeof(vals));
    return 0;
  } catch (...) {
    return -1;
  }
}

int sankhya_pdhg_memory_check(const sankhya_model* h, int precision, int64_t free_bytes,
                              int64_t total_bytes, int64_t* required, int64_t* reserve) {
  if (!h) return -1;
  try {
    const Model& md = h->model;
    const hip_engine::Prepared prep = hip_engine::prepare_unscaled(md, 1);
    const int64_t need = pdhg_dev_estimate_bytes(
        md.num_rows(), md.num_cols(), md.num_nonzeros(), prep.a_rp.data(), prep.a_ci.data(),
        md.matrix.column_starts().data(), md.matrix.row_indices().data(), precision, 1, 0);
    if (need < 0) return -1;
    const hip_engine::MemoryCheck mc = hip_engine::check_device_memory(need, free_bytes, total_bytes);
    if (required) *required = need;
    if (reserve) *reserve = pdhg_dev_vram_reserve(total_bytes);
    return mc.fits ? 1 : 0;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
