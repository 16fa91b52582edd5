// SPDX-License-Identifier: Apache-2.0
// Host steps shared by the single-LP and the batched engine: the problem's norms and
// preconditioning with the CSR the device reads (src/pdhg/pdhg.cpp:94-117), the power
// iteration's start vector (src/la/scaling.cpp:168-171), and the report that turns the device's
// final residuals into a Solution (src/pdhg/pdhg.cpp:486-567).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "sankhya/model.hpp"
#include "sankhya/options.hpp"
#include "sankhya/logging.hpp"
#include "scaling.hpp"

namespace sankhya::hip_engine {

constexpr int kRuizIterations = 10;   // src/pdhg/pdhg.cpp:62
constexpr int kPowerIterations = 30;  // src/pdhg/pdhg.cpp:63

struct Residuals {
  double primal = 0, dual = 0, gap = 0, primal_objective = 0, dual_objective = 0;
  double absolute_primal = 0, absolute_dual = 0, gap_as_verified = 0, complementarity = 0;
  // src/pdhg/pdhg_evaluate.hpp:57-60.
  [[nodiscard]] bool meets_request(double tolerance) const {
    return primal <= tolerance && dual <= tolerance && gap <= tolerance &&
           absolute_primal <= tol::kPrimalFeasibility;
  }
  [[nodiscard]] bool meets_standard() const {
    return absolute_primal <= tol::kPrimalFeasibility && absolute_dual <= tol::kDualFeasibility &&
           gap_as_verified <= tol::kDualityGap && complementarity <= 1e-6;
  }
};
Residuals from_array(const double* r);

// Storage whose resize leaves new elements of a trivial type uninitialised, for arrays written
// whole right afterwards: zero-filling the fresh 60 MB of a 5M-nonzero CSR on one thread (the
// first touch of every page) took longer than the parallel transpose that then fills it.
template <class T>
struct UninitAllocator : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAllocator<U>;
  };
  UninitAllocator() noexcept = default;
  template <class U>
  UninitAllocator(const UninitAllocator<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible_v<U>) {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using HostArray = std::vector<T, UninitAllocator<T>>;

// CSR of a CSC matrix, columns ascending within each row (the counting-sort transpose of
// src/la/sparse.cpp:271-302); threads > 1 gives the same arrays (matrices under 200,000
// nonzeros use one thread).
void csc_to_csr(Index m, Index n, const std::vector<Index>& cp, const std::vector<Index>& ri,
                const std::vector<double>& v, std::vector<Index>* rp, HostArray<Index>* ci,
                HostArray<double>* cv, int threads = 1);

// One LP as the device receives it.
struct Prepared {
  std::vector<double> cost;  // minimise-space cost
  double cost_norm = 0.0, bound_norm = 0.0;
  Scaling scaling;
  std::vector<Index> a_rp;  // CSR of Ahat
  HostArray<Index> a_ci;
  HostArray<double> a_v;
  double scaling_seconds = 0.0, csr_seconds = 0.0;  // host time of the two steps
};
// threads: host threads for the preconditioner (the batch prepares LPs in parallel with 1).
Prepared prepare(const Model& model, int threads = 1);
// The same without the host preconditioner: `scaling` stays empty and a_v holds A's own values,
// for pdhg_dev_upload_unscaled, which scales on the device.
Prepared prepare_unscaled(const Model& model, int threads = 1);

std::vector<double> power_start(const Model& model, std::int64_t random_seed);

struct ReportState {
  bool converged;
  SolveStatus stop_status;
  std::int64_t iterations, restarts;
  double tolerance;
  bool stop_at_request;
  bool relative_kkt;  // the device stopped under pdhg_stop_rule=relative_kkt
};
// Duals with the reference's sign flip, status and message, dual bound, recompute_quality.
// col_value must already hold the point; solve_seconds is left to the caller. With `prep` (from
// prepare_unscaled: the CSR of A's own values) the quality passes run on the host's threads.
void report(const Model& model, const double* res9, const std::vector<double>& y,
            const std::vector<double>& reduced, const ReportState& rs, Solution* solution,
            const Prepared* prep = nullptr);

// Solution::recompute_quality (src/core/model.cpp:247-441) with its passes over the nonzeros
// spread over `threads` host threads: the row activities and row scales walk the CSR of A's own
// values (rp, ci, cv) by rows, each row's terms in ascending column order, the order in which
// the reference's column-by-column scatter adds them; the implied reduced costs walk the
// columns; the passes that take maxima over rows or columns stay serial, in the reference's
// order. The quality fields equal the serial function's bit for bit.
void recompute_quality_threaded(const Model& model, const std::vector<Index>& rp,
                                const HostArray<Index>& ci, const HostArray<double>& cv,
                                int threads, Solution* s);

// The dispatcher's VRAM decision (src/core/solve.cpp:751-771): the estimate plus the reserve
// (max(256 MiB, total / 10), src/gpu/gpu_memory.cpp:50-55) must fit in the free bytes. Host
// arithmetic only.
struct MemoryCheck {
  bool fits;
  double required_mib, free_mib, reserve_mib;
  std::string message;  // set when it does not fit
};
MemoryCheck check_device_memory(std::int64_t required, std::int64_t free_bytes,
                                std::int64_t total_bytes);

// The device the gpu_devices option names (one GPU: the first listed, with a warning when
// several are given).
// Host threads for one LP's preparation: the machine's, at most 16.
int host_threads();

int engine_device(const Options& options, Logger& logger);

// prepare() (scale_on_host) or prepare_unscaled() and power_start() for every model, on all
// host cores (one model per task).
void prepare_all(const std::vector<const Model*>& models, std::int64_t random_seed,
                 std::vector<Prepared>* prep, std::vector<std::vector<double>>* v0,
                 bool scale_on_host = true);

// A batch's arrays, concatenated LP after LP in the layout include/pdhg_batch.h takes. From
// prepare_unscaled() results the values are A's own and the scaled vectors and Dr/Dc are left
// empty (pdhg_batch_upload_unscaled).
struct BatchInput {
  std::vector<int32_t> m, n;
  std::vector<int64_t> nnz;
  std::vector<int32_t> a_rp, a_ci, t_rp, t_ci;
  std::vector<double> a_v, t_v, chat, lhat, uhat, rlhat, ruhat, dc, dr, cost, cl, cu, rl, ru;
  std::vector<double> bound_norm, cost_norm, v0;
};
BatchInput make_batch_input(const std::vector<const Model*>& models,
                            const std::vector<Prepared>& prep,
                            const std::vector<std::vector<double>>& v0, bool host_scaled = true);

}  // namespace sankhya::hip_engine
