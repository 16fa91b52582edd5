// SPDX-License-Identifier: Apache-2.0
// Diagonal preconditioning of the LP, Ahat = Dr A Dc (reference: src/la/scaling.hpp:36-71,
// src/la/scaling.cpp:14-159): 10 Ruiz infinity-norm rounds, then one Pock-Chambolle round with
// alpha = 1. Values are scaled in place in the model's CSC order, each entry as
// (a * r_i) * c_j per pass, so the scaled matrix and multipliers equal the reference's bit for
// bit (the reference rebuilds the matrix each pass; its pattern and order do not change).
#pragma once

#include <vector>

#include "sankhya/model.hpp"

namespace sankhya::hip_engine {

struct Scaling {
  std::vector<double> row, column;    // Dr, Dc
  std::vector<double> values;         // Ahat in the model's CSC order
  std::vector<double> cost, col_lower, col_upper, row_lower, row_upper;
  double max_abs = 0.0, min_abs = 0.0;
};

// threads > 1 splits the passes over that many host threads with the same result bit for
// bit (models under 200,000 nonzeros always use one).
[[nodiscard]] Scaling build_scaling(const Model& model, const std::vector<double>& min_space_cost,
                                    int ruiz_iterations, int threads = 1);

}  // namespace sankhya::hip_engine
