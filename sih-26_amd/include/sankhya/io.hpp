// SPDX-License-Identifier: Apache-2.0
// Model input and solution output (reference: include/sankhya/io.hpp).
//
// Reading (reference include/sankhya/io.hpp:19-53): MPS in its free and fixed dialects. A read
// returns a status, never throws: a malformed file is ordinary input, and every message names
// the file and the 1-based line. A reader validates the model before it reports success.
// CPLEX LP files and quadratic (QPS) sections are refused with a message naming the reason:
// the engine here solves linear programs. Gzip-compressed files are read through zlib.
//
// The solution file, the engine's certificate output (reference: include/sankhya/io.hpp:56-72,
// format of src/io/writer.cpp:121-284):
// A line-oriented text file that an independent checker re-reads against the original model:
// a header of `key value` lines, then `begin columns n` ... `end columns` (name, value, reduced
// cost, basis status) and `begin rows m` ... `end rows` (name, activity, dual, basis status).
// Every number is written with 17 significant digits, so each double reads back bit for bit;
// infinities are the words `inf` and `-inf`; a negative zero is written as 0. A name holding
// whitespace, a quote or a backslash is written in double quotes with `"` and `\` escaped.
//
// PDHG proves no infeasibility and finds no ray, so the header always says `certificate none`.
// A status that claims no point (model_error, numerical_error, not_solved) writes the header
// alone, without the columns and rows blocks.
#pragma once

#include <string>
#include <utility>

#include "sankhya/model.hpp"
#include "sankhya/options.hpp"

namespace sankhya::io {

/// Outcome of a read. When `ok` is false the model's contents are unspecified.
struct ReadResult {
  bool ok = false;
  std::string error;  ///< "path:line: message"; empty when ok

  [[nodiscard]] explicit operator bool() const noexcept { return ok; }
  static ReadResult success() { return ReadResult{true, {}}; }
  static ReadResult failure(std::string message) { return ReadResult{false, std::move(message)}; }
};

/// MPS dialect. kAuto reads free MPS and falls back to fixed columns when the free reading
/// fails, which only happens for fixed files whose names hold blanks.
enum class MpsFormat { kAuto, kFree, kFixed };

/// "auto", "free" or "fixed" (any case) as an MpsFormat; false for anything else.
[[nodiscard]] bool parse_mps_format(const std::string& text, MpsFormat* out) noexcept;

/// Read an MPS file into `model`, replacing it. `format_used`, when non-null, receives the
/// dialect that read the file.
ReadResult read_mps(const std::string& path, Model* model, MpsFormat format = MpsFormat::kAuto,
                    MpsFormat* format_used = nullptr);

/// Read a model by its extension: `.lp` (with or without `.gz`) is refused, anything else is
/// read as MPS.
ReadResult read_model(const std::string& path, Model* model);

/// Write `solution` for `model` to `path`. Returns false and fills `error` on an I/O failure.
/// The header records the default MIP gap targets.
bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    std::string* error);

/// The same, recording the `mip_relative_gap` and `mip_absolute_gap` of `options`.
bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    const Options& options, std::string* error);

}  // namespace sankhya::io
