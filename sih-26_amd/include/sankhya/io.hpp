// SPDX-License-Identifier: Apache-2.0
// The solution file, the engine's certificate output (reference: include/sankhya/io.hpp:56-72,
// format of src/io/writer.cpp:121-284).
//
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

#include "sankhya/model.hpp"
#include "sankhya/options.hpp"

namespace sankhya::io {

/// Write `solution` for `model` to `path`. Returns false and fills `error` on an I/O failure.
/// The header records the default MIP gap targets.
bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    std::string* error);

/// The same, recording the `mip_relative_gap` and `mip_absolute_gap` of `options`.
bool write_solution(const std::string& path, const Model& model, const Solution& solution,
                    const Options& options, std::string* error);

}  // namespace sankhya::io
