// SPDX-License-Identifier: Apache-2.0
// The model handle shared by the two C ABIs (include/sankhya_pdhg.h, include/sankhya_c.h).
// Reference counterpart: struct sankhya_model in src/api/c_api.cpp:108-119.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <utility>

#include "sankhya/model.hpp"
#include "sankhya/solve_control.hpp"

struct sankhya_model {
  sankhya::Model model;
  // Coefficients set one by one (sankhya_model_set_coefficient), pending until a solve builds
  // the matrix; `pending` is false while `model.matrix` is the whole truth.
  std::map<std::pair<int, int>, double> entries;
  bool pending = false;
  // The SolveControl of the solve in progress, for sankhya_model_interrupt from other threads.
  std::mutex control_mutex;
  std::shared_ptr<sankhya::SolveControl> active_control;
  sankhya::ProgressCallback progress_callback;
};
