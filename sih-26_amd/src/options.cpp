// SPDX-License-Identifier: Apache-2.0
// Option registry rows the PDHG path and its dispatcher read (reference:
// src/util/options.cpp:83-800; defaults as there), plus this engine's own rows.
#include "sankhya/options.hpp"

#include <algorithm>
#include <cassert>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <sstream>

#include "sankhya/tolerances.hpp"

namespace sankhya {
namespace {

constexpr double kNoLimit = std::numeric_limits<double>::infinity();
constexpr double kBig = std::numeric_limits<double>::max();

std::vector<OptionSpec> build_registry() {
  std::vector<OptionSpec> s;
  s.push_back({"time_limit", OptionType::Double, kNoLimit, "Wall-clock limit in seconds.", 0.0, kNoLimit});
  s.push_back({"iteration_limit", OptionType::Int, std::int64_t{-1},
               "PDHG iteration limit; -1 for the engine's own ceiling of 1e6.", -1.0, 9.0e18});
  s.push_back({"algorithm", OptionType::String, std::string("auto"), "LP engine: auto or pdhg.", 0, 0,
               {"auto", "simplex", "dual-simplex", "pdhg", "ipm"}});
  s.push_back({"presolve", OptionType::Bool, true,
               "Presolve reductions; not part of this engine, which solves the model as given.", 0, 0,
               {}, "out of scope"});
  s.push_back({"gpu", OptionType::Bool, false,
               "Select the GPU PDHG backend. This build has only the MI355X engine, so PDHG runs "
               "on the GPU either way.", 0, 0});
  s.push_back({"gpu_devices", OptionType::String, std::string("auto"),
               "HIP device ID to run on; 'auto' is device 0. A list uses its first entry.", 0, 0, {}, "",
               true});
  s.push_back({"threads", OptionType::Int, std::int64_t{1},
               "Host worker threads of the reference's CPU engine; accepted and not read here (the host "
               "preparation of one large LP uses up to 16 threads).", 0.0, 1024.0});
  s.push_back({"deterministic", OptionType::Bool, false,
               "Reproducible solves. The MI355X engine uses no floating-point atomics, so its "
               "results repeat bit for bit and it is admitted under this option; the batched "
               "engine then skips its timing-based schedule tuner.", 0, 0});
  s.push_back({"random_seed", OptionType::Int, std::int64_t{0}, "Seed; PDHG's power iteration uses seed + 1.",
               0.0, 4294967295.0});
  s.push_back({"pdhg_restart", OptionType::Bool, true, "Restart PDHG on the KKT-error criterion.", 0, 0});
  s.push_back({"pdhg_polish", OptionType::Bool, true,
               "Hand a PDHG point that is not optimal to the interior-point polish as a warm "
               "start (col_value, row_dual, col_dual). solve() runs it when an interior-point "
               "engine is installed (sankhya::install_polish_engine, sankhya/ipm.hpp); the PDHG "
               "engine itself never reads this option.", 0, 0});
  s.push_back({"pdhg_mixed_polish", OptionType::Bool, true,
               "Mixed precision: after the FP32 phase, continue the same run in FP64 on the GPU "
               "to pdhg_tolerance (the FP64 polish). Off: the FP32 phase runs to the end.", 0, 0});
  s.push_back({"polish_iteration_limit", OptionType::Int, std::int64_t{50},
               "Interior-point iterations the polish may spend.", 1.0, 9.0e18});
  s.push_back({"polish_max_seconds", OptionType::Double, 30.0,
               "Wall-clock the polish may spend: the smaller of this and what time_limit has "
               "left.", 0.1, kNoLimit});
  s.push_back({"polish_start_margin", OptionType::Double, 1e-1,
               "Read by the installed interior-point engine: the floor of its warm start's "
               "slacks and multipliers.", 1e-8, 1.0});
  s.push_back({"polish_max_factor_nonzeros", OptionType::Int, std::int64_t{50000000},
               "Read by the installed interior-point engine: it declines the polish above this "
               "many factor nonzeros; -1 for no cap.", -1.0, 9.0e18});
  s.push_back({"pdhg_tolerance", OptionType::Double, tol::kPdhgLoose, "PDHG relative KKT tolerance.", 0.0, 1.0});
  s.push_back({"pdhg_stop_at_request", OptionType::Bool, false,
               "Stop once pdhg_tolerance is met and the point is primal-feasible to the absolute "
               "tolerance, without running on to the project standard.", 0, 0});
  s.push_back({"pdhg_precision", OptionType::String, std::string("fp64"),
               "Bulk arithmetic: fp64, or mixed (FP32 matrix and iterates, FP64 reductions, step "
               "control and residuals).", 0, 0, {"fp64", "mixed"}});
  s.push_back({"pdhg_mixed_switch", OptionType::Double, 1e-6,
               "Mixed precision with pdhg_mixed_polish: the FP32 phase runs until the relative KKT error "
               "is <= max(this, pdhg_tolerance), or stalls; the same run then continues in FP64 "
               "to pdhg_tolerance.", 0.0, 1.0});
  s.push_back({"pdhg_stop_rule", OptionType::String, std::string("reference"),
               "reference: stop as SANKHYA's PDHG does (relative residuals <= pdhg_tolerance and "
               "absolute primal <= 1e-7, and the project standard unless pdhg_stop_at_request). "
               "relative_kkt: stop when the relative KKT error, max(relative primal, dual, gap), "
               "is <= pdhg_tolerance; the report still only claims optimal on the standard.",
               0, 0, {"reference", "relative_kkt"}});
  s.push_back({"pdhg_batch_tune", OptionType::Bool, true,
               "Batched engine: time the three workgroup geometries on every LP and keep the "
               "fastest per LP before solving.", 0, 0});
  s.push_back({"pdhg_batch_tune_attempts", OptionType::Int, std::int64_t{96},
               "Batched engine: attempts per geometry when tuning the schedule.", 1.0, 100000.0});
  s.push_back({"mip_relative_gap", OptionType::Double, tol::kMipRelativeGap,
               "MIP relative gap target; this engine solves LPs and only records it in solution "
               "files.", 0.0, 1.0});
  s.push_back({"mip_absolute_gap", OptionType::Double, tol::kMipAbsoluteGap,
               "MIP absolute gap target; recorded in solution files.", 0.0, kNoLimit});
  s.push_back({"pdhg_chunk", OptionType::Int, std::int64_t{0},
               "Attempts per captured graph chunk; 0 picks the default.", 0.0, 4096.0});
  s.push_back({"log_level", OptionType::String, std::string("info"), "off, error, warning, info, verbose, debug.",
               0, 0, {"off", "error", "warning", "info", "verbose", "debug"}});
  s.push_back({"log_to_console", OptionType::Bool, true, "Write the solver log to stdout.", 0, 0});
  s.push_back({"progress_out", OptionType::String, std::string(""),
               "Append live progress as JSON lines to this file.", 0, 0, {}, "", true});
  s.push_back({"profile", OptionType::String, std::string("off"), "off, basic, detailed.", 0, 0,
               {"off", "basic", "detailed"}});
  s.push_back({"primal_feasibility_tolerance", OptionType::Double, tol::kPrimalFeasibility,
               "Max allowed row/column bound violation.", 0.0, kBig});
  s.push_back({"dual_feasibility_tolerance", OptionType::Double, tol::kDualFeasibility,
               "Max allowed reduced-cost sign violation.", 0.0, kBig});
  return s;
}

std::string lower(std::string v) {
  for (char& c : v) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return v;
}

}  // namespace

const std::vector<OptionSpec>& Options::registry() {
  static const std::vector<OptionSpec> r = build_registry();
  return r;
}

const OptionSpec* Options::find_spec(const std::string& name) {
  for (const OptionSpec& s : registry())
    if (s.name == name) return &s;
  return nullptr;
}

bool Options::exists(const std::string& name) { return find_spec(name) != nullptr; }

Options::Options() {
  for (const OptionSpec& s : registry()) values_.push_back(s.default_value);
}

std::size_t Options::index_of(const std::string& name) const {
  const auto& r = registry();
  for (std::size_t i = 0; i < r.size(); ++i)
    if (r[i].name == name) return i;
  assert(false && "unknown option");
  std::abort();
}

void Options::set_bool(const std::string& name, bool value) { values_[index_of(name)] = value; }
void Options::set_int(const std::string& name, std::int64_t value) { values_[index_of(name)] = value; }
void Options::set_double(const std::string& name, double value) { values_[index_of(name)] = value; }
void Options::set_string(const std::string& name, const std::string& value) {
  const std::size_t i = index_of(name);
  values_[i] = registry()[i].case_sensitive ? value : lower(value);
}
bool Options::get_bool(const std::string& name) const { return std::get<bool>(values_[index_of(name)]); }
std::int64_t Options::get_int(const std::string& name) const {
  return std::get<std::int64_t>(values_[index_of(name)]);
}
double Options::get_double(const std::string& name) const { return std::get<double>(values_[index_of(name)]); }
const std::string& Options::get_string(const std::string& name) const {
  return std::get<std::string>(values_[index_of(name)]);
}

bool Options::is_modified(const std::string& name) const {
  const std::size_t i = index_of(name);
  return values_[i] != registry()[i].default_value;
}

std::vector<std::string> Options::modified_names() const {
  std::vector<std::string> out;
  for (std::size_t i = 0; i < values_.size(); ++i)
    if (values_[i] != registry()[i].default_value) out.push_back(registry()[i].name);
  return out;
}

std::string Options::value_as_string(const std::string& name) const {
  const OptionValue& v = values_[index_of(name)];
  std::ostringstream o;
  if (std::holds_alternative<bool>(v)) o << (std::get<bool>(v) ? "true" : "false");
  else if (std::holds_alternative<std::int64_t>(v)) o << std::get<std::int64_t>(v);
  else if (std::holds_alternative<double>(v)) {
    o.precision(17);
    o << std::get<double>(v);
  } else o << std::get<std::string>(v);
  return o.str();
}

bool Options::set_from_string(const std::string& name, const std::string& text, std::string* error) {
  const OptionSpec* spec = find_spec(name);
  auto bad = [&](const std::string& why) {
    if (error) *error = name + ": " + why;
    return false;
  };
  if (!spec) return bad("unknown option");
  const std::size_t i = index_of(name);
  switch (spec->type) {
    case OptionType::Bool: {
      const std::string t = lower(text);
      if (t == "true" || t == "1" || t == "on" || t == "yes") values_[i] = true;
      else if (t == "false" || t == "0" || t == "off" || t == "no") values_[i] = false;
      else return bad("not a boolean: " + text);
      return true;
    }
    case OptionType::Int: {
      char* end = nullptr;
      const long long v = std::strtoll(text.c_str(), &end, 10);
      if (text.empty() || *end != '\0') return bad("not an integer: " + text);
      if (static_cast<double>(v) < spec->min_value || static_cast<double>(v) > spec->max_value)
        return bad("out of range: " + text);
      values_[i] = static_cast<std::int64_t>(v);
      return true;
    }
    case OptionType::Double: {
      char* end = nullptr;
      const double v = std::strtod(text.c_str(), &end);
      if (text.empty() || *end != '\0' || std::isnan(v)) return bad("not a number: " + text);
      if (v < spec->min_value || v > spec->max_value) return bad("out of range: " + text);
      values_[i] = v;
      return true;
    }
    case OptionType::String: {
      const std::string v = spec->case_sensitive ? text : lower(text);
      if (!spec->choices.empty() &&
          std::find(spec->choices.begin(), spec->choices.end(), v) == spec->choices.end())
        return bad("not one of the permitted values: " + text);
      values_[i] = v;
      return true;
    }
  }
  return bad("unhandled type");
}

}  // namespace sankhya
