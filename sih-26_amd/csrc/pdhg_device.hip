// pdhg_device.hip - MI355X (gfx950) restarted-PDHG engine: kernels, device state, graph
// capture and the C ABI declared in include/pdhg_hip.h.
//
// One PDHG attempt (src/pdhg/pdhg.cpp:220-335) is three kernels:
//   k_primal  Ahat^T y gathered row by row over the stored transpose, fused with the primal
//             step x' = proj(x - tau (c + Ahat^T y)) and the primal movement term
//             (pdhg.cpp:225-236,252-256; the CUDA path's cusparseSpMV(T) + k_primal_fused,
//             src/gpu/pdhg_gpu.cu:90-112,566-578).
//   k_dual    Ahat [ext, dx] for both right-hand sides in ONE pass over Ahat, fused with the
//             Moreau-prox dual step, the dual movement and the interaction term
//             (pdhg.cpp:238-277; the CUDA path reads Ahat twice, pdhg_gpu.cu:581-608).
//   k_step    one workgroup: adds the partial sums in index order and applies the adaptive
//             step rule, acceptance and the evaluation-tick test on the device
//             (pdhg.cpp:293-342), so no host round trip happens per attempt.
// The iterates are stored as interleaved pairs (x_a, x_b) so that a single 16-byte gather
// in k_dual yields both x and x' and therefore ext = 2x'-x and dx = x'-x.
// Acceptance flips which half of each pair is current; the running sums are updated lazily
// by the next attempt's kernels (or at the tick), so an accept costs no extra pass.
//
// A graph chunk is `chunk` attempt slots followed by the evaluation tick (pdhg.cpp:341-483):
//   k_eval_rows / k_eval_cols   the convergence test of both candidates (current and average)
//                               in one pass over Ahat and one over Ahat^T,
//   k_decide                    residuals, best tracking, stopping test, restart and primal
//                               weight (one workgroup),
//   k_apply / k_clear           copies the best point, resets or advances the running sums.
// Slots after a pending tick, and the tick kernels when none is due, exit at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "pdhg_iterate.hpp"
#include "pdhg_kernels.hpp"
#include "plan_host.hpp"
#include "scaling_waves.hpp"

namespace sankhya_hip {

// Attempt slots per graph chunk: the evaluation interval and two spare slots. A tick is due
// after kEvalInterval accepted attempts, and nearly every attempt is accepted (20,512 attempts
// for 20,400 iterations on the 200k x 500k Zipf LP), so the tick usually falls in the chunk's
// last three slots. Every slot after it still launches its three kernels, which exit at once
// (about 7.5 us a slot). A cycle that needs more attempts than the chunk holds spills into a
// second chunk. Measured per iteration: 42 slots 56.7 us, 40 57.6, 48 57.8, 64 60.8.
constexpr int kDefaultChunk = static_cast<int>(kEvalInterval) + 2;

constexpr size_t kBlockBytes = size_t{64} << 20;  // smallest device block a handle allocates

// Workgroup sizes of the two iteration kernels (tiles as in SingleGeo).
constexpr int kPrimalThreads = 256;
constexpr int kDualThreads = 256;
using PrimalGeo = Geometry<kPrimalThreads, kTile, kSeg, kMaxRows>;
using DualGeo = Geometry<kDualThreads, kTile, kSeg, kMaxRows>;


// ---- iteration kernels ------------------------------------------------------------------

// NT: the matrix loads' flavour (non-temporal, or cached when the LP fits the MALL).
template <typename V, typename X, bool NT>
__global__ __launch_bounds__(kPrimalThreads) void k_primal(PlanView<V> T, Iter<X> it, Ctl* ctl,
                                                           double* part, int force) {
  __shared__ double lds[kTile];
  __shared__ double red[(kPrimalThreads / 64) * 2];
  const int4 item = T.items[blockIdx.x];
  const AttemptCtl c = read_attempt(ctl);
  if (attempt_halted(c, force, item,
                     iter_bits(it) | plan_bits(T) | reinterpret_cast<uintptr_t>(part)))
    return;
  PrimalGather<X> gather{it.ybuf, c.cur};
  PrimalEpi<X> epi = primal_epi(it, c);
  Acc<1, 0> acc;
  acc.zero();
  run_item_geo<PrimalGeo, V, 1, NT>(T, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc.s[0];
}

// Long columns of Ahat (dense columns), completed before k_dual reads x'.
template <typename V, typename X>
__global__ __launch_bounds__(kTail) void k_primal_tail(PlanView<V> T, Iter<X> it, Ctl* ctl,
                                                       double* lr_sum, int force) {
  __shared__ double lds[kChunk];
  __shared__ double red[16 * 1];
  // The long columns' inputs were staged by k_primal: the halt test waits on no other load.
  const AttemptCtl c = read_attempt(ctl);
  const int halt = force ? 0 : c.halt;
  PrimalEpi<X> epi = primal_epi(it, c);
  Acc<1, 0> acc;
  acc.zero();
  if (!finish_staged_rows<V, 1>(T, epi, acc, lds, [&]() { return halt == 0; })) return;
  block_reduce(acc, red);
  if (threadIdx.x == 0) lr_sum[0] = acc.s[0];
}

// stepf (2 doubles): the step rule's factors for this attempt, evaluated by workgroup 0 while the
// others work, and read by k_dual_tail.
template <typename V, typename X, bool NT>
__global__ __launch_bounds__(kDualThreads) void k_dual(PlanView<V> A, Iter<X> it, Ctl* ctl,
                                                       double* part, double* stepf, int force) {
  __shared__ double lds[kTile];
  __shared__ double red[(kDualThreads / 64) * 2];
  const int4 item = A.items[blockIdx.x];
  const AttemptCtl c = read_attempt(ctl);
  if (attempt_halted(c, force, item,
                     iter_bits(it) | plan_bits(A) | reinterpret_cast<uintptr_t>(part)))
    return;
  DualGather<X> gather{it.xbuf, c.cur};
  DualEpi<X> epi = dual_epi(it, c);
  Acc<2, 0> acc;
  acc.zero();
  run_item_geo<DualGeo, V, 2, NT>(A, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc.s[0];
    part[2 * blockIdx.x + 1] = acc.s[1];
  }
  // After its own item, so that the iteration count's load and the two pow calls do not
  // delay the item's loads.
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const StepFactors f = step_factors(ctl->iteration);
    stepf[0] = f.shrink;
    stepf[1] = f.grow;
  }
}

// Completes the long rows of Ahat, then adds every partial in index order and takes the
// step decision: one workgroup, after k_primal(+tail) and k_dual.
template <typename V, typename X>
__global__ __launch_bounds__(kTail) void k_dual_tail(PlanView<V> A, Iter<X> it, Ctl* ctl,
                                                     const double* part_p,
                                                     const double* lr_p, int has_lr_p,
                                                     const double* part_d, int nd, int rounds,
                                                     const double* stepf, int force) {
  __shared__ double lds[2 * kChunk];
  __shared__ double red[16 * 5];
  // Every load of this kernel is independent of every other (the long rows' inputs were
  // staged by k_dual), so all of them, the halt flag included, share one memory round trip:
  // the halt test comes after they are issued and before anything is written. The control
  // fields are read by every thread (uniform scalar loads, no branch to wait at).
  const AttemptCtl c = read_attempt(ctl);
  const int halt = force ? 0 : c.halt;
  DualEpi<X> epi = dual_epi(it, c);
  const StepSnap snap = step_snap(ctl);
  // Workgroup 0 of k_dual wrote them; with no dual items they are evaluated here.
  const double f0 = stepf[0], f1 = stepf[1];
  // s[0] primal movement, s[1] dual movement, s[2] interaction (per-workgroup partials);
  // s[3], s[4] the same two dual terms from the long rows.
  Acc<5, 0> a;
  a.zero();
  auto partials = [&]() {
    double p[1] = {0.0}, d[2] = {0.0, 0.0};
    lane_sums_1_2_padded(part_p, part_d, rounds, p, d);
    if (has_lr_p && threadIdx.x == 0) p[0] += lr_p[0];
    a.s[0] = p[0];
    a.s[1] = d[0];
    a.s[2] = d[1];
    return halt == 0;
  };
  Acc<2, 0> lacc;
  lacc.zero();
  if (!finish_staged_rows<V, 2>(A, epi, lacc, lds, partials)) return;
  a.s[3] = lacc.s[0];
  a.s[4] = lacc.s[1];
  block_reduce(a, red);
  if (threadIdx.x == 0) {
    const StepFactors f = nd > 0 ? StepFactors{f0, f1} : step_factors(snap.iteration);
    step_decide(ctl, snap, a.s[0], a.s[1] + a.s[3], a.s[2] + a.s[4], f);
  }
}

template <typename V, typename X, bool NT>
__global__ __launch_bounds__(kThreads) void k_eval_rows(PlanView<V> A, Iter<X> it, Ctl* ctl,
                                                        double* part) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 9];
  const int4 item = A.items[blockIdx.x];
  if (!ctl->tick) return;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickRowsGather<X> gather{it, cur, tp, count};
  TickRowsEpi<X> epi{it, cur, tp, count};
  Acc<7, 2> acc;
  acc.zero();
  run_item<V, 2, NT>(A, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) part[9 * blockIdx.x + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) part[9 * blockIdx.x + 7 + k] = acc.m[k];
  }
}

template <typename V, typename X, bool NT>
__global__ __launch_bounds__(kThreads) void k_eval_cols(PlanView<V> T, Iter<X> it, Ctl* ctl,
                                                        double* part) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 9];
  const int4 item = T.items[blockIdx.x];
  if (!ctl->tick) return;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickColsGather<X> gather{it, cur, tp, count};
  TickColsEpi<X> epi{it, cur, tp, count};
  Acc<7, 2> acc;
  acc.zero();
  run_item<V, 2, NT>(T, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) part[9 * blockIdx.x + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) part[9 * blockIdx.x + 7 + k] = acc.m[k];
  }
}

// Best tracking, stopping test, restart and primal weight (src/pdhg/pdhg.cpp:344-483), after
// completing the long rows of both products.
template <typename V, typename X>
__global__ __launch_bounds__(kTail) void k_decide(PlanView<V> A, PlanView<V> T, Iter<X> it,
                                                  Ctl* ctl, const double* per, const double* pec,
                                                  int rounds, pdhg_tick* ticks) {
  __shared__ double lds[2 * kChunk];
  __shared__ double red[16 * 9];
  if (!ctl->tick) return;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickRowsEpi<X> repi{it, cur, tp, count};
  TickColsEpi<X> cepi{it, cur, tp, count};
  Acc<7, 2> ra, ca;
  ra.zero();
  ca.zero();
  // The long rows' and long columns' inputs were staged by k_eval_rows and k_eval_cols.
  finish_staged_rows<V, 2>(A, repi, ra, lds, NoPre());
  finish_staged_rows<V, 2>(T, cepi, ca, lds, NoPre());
  block_reduce(ra, red);
  block_reduce(ca, red);
  constexpr bool kMax9[9] = {false, false, false, false, false, false, false, true, true};
  double vr[9], vc[9];
  // part_er and part_ec hold `rounds` whole rounds of partials, zero past the items.
  block_sums2_padded<9>(per, pec, rounds, 9, kMax9, vr, vc, red);
  if (threadIdx.x != 0) return;
  merge9(vr, ra);
  merge9(vc, ca);
  tick_decide(ctl, vr, vc, ticks);
}

template <typename X>
__global__ __launch_bounds__(kThreads) void k_apply(Iter<X> it, Ctl* ctl) {
  if (!ctl->apply) return;
  const ApplyArgs args = apply_args(ctl);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t j = blockIdx.x * (int64_t)kThreads + threadIdx.x; j < it.n; j += stride)
    apply_col(it, args, j);
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < it.m; i += stride)
    apply_row(it, args, i);
}

// Exact FP32 -> FP64 widening of the iterate pairs (the mixed-precision handoff).
__global__ void k_widen(const float* xs, const double* xr, double* xd, int64_t nx,
                        const float* ys, const double* yr, double* yd, int64_t ny) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nx; k += stride)
    xd[k] = xr[k >> 1] + static_cast<double>(xs[k]);
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ny; k += stride)
    yd[k] = yr[k >> 1] + static_cast<double>(ys[k]);
}

__global__ void k_clear(Ctl* ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->apply = 0;
    ctl->copy_best = 0;
    ctl->restart_now = 0;
  }
}

// ---- start, report, test hooks ----------------------------------------------------------

template <typename X>
__global__ void k_init(Iter<X> it) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < it.n; j += stride)
    init_col(it, j);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < it.m; i += stride)
    init_row(it, i);
}

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_single_rows(PlanView<V> A, Iter<X> it,
                                                          const double* bx, const double* by,
                                                          double* act_out, double* part) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 4];
  const int4 item = A.items[blockIdx.x];
  SingleRowsGather<X> gather{it, bx};
  SingleRowsEpi<X> epi{it, by, act_out};
  Acc<3, 1> acc;
  acc.zero();
  run_item<V, 1>(A, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) part[4 * blockIdx.x + k] = acc.s[k];
    part[4 * blockIdx.x + 3] = acc.m[0];
  }
}

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_single_cols(PlanView<V> T, Iter<X> it,
                                                          const double* bx, const double* by,
                                                          double* red_out, double* part) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 4];
  const int4 item = T.items[blockIdx.x];
  SingleColsGather<X> gather{it, by};
  SingleColsEpi<X> epi{it, bx, red_out};
  Acc<3, 1> acc;
  acc.zero();
  run_item<V, 1>(T, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) part[4 * blockIdx.x + k] = acc.s[k];
    part[4 * blockIdx.x + 3] = acc.m[0];
  }
}

template <typename V, typename X>
__global__ __launch_bounds__(kTail) void k_single_decide(PlanView<V> A, PlanView<V> T, Iter<X> it,
                                                         Ctl* ctl, const double* bx,
                                                         const double* by, double* act_out,
                                                         double* red_out, const double* pr,
                                                         int npr, const double* pc, int npc,
                                                         double* res9) {
  __shared__ double lds[kChunk];
  __shared__ double red[16 * 4];
  SingleRowsEpi<X> repi{it, by, act_out};
  SingleColsEpi<X> cepi{it, bx, red_out};
  Acc<3, 1> ra, ca;
  ra.zero();
  ca.zero();
  finish_long_rows<V, 1>(A, repi, ra, lds);
  finish_long_rows<V, 1>(T, cepi, ca, lds);
  block_reduce(ra, red);
  block_reduce(ca, red);
  constexpr bool kMax4[4] = {false, false, false, true};
  double vr[4], vc[4];
  block_sums<4>(pr, npr, nullptr, 0, 4, kMax4, vr, red);
  block_sums<4>(pc, npc, nullptr, 0, 4, kMax4, vc, red);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 3; ++k) {
    vr[k] += ra.s[k];
    vc[k] += ca.s[k];
  }
  vr[3] = smax(vr[3], ra.m[0]);
  vc[3] = smax(vc[3], ca.m[0]);
  // rows: 0 pv 1 sup 2 dv 3 cm ; cols: 0 dv 1 bc 2 po 3 cm
  const Res r = make_res(vr[0], vc[0] + vr[2], smax(vr[3], vc[3]), vc[2], vc[1] - vr[1],
                         ctl->bound_norm, ctl->cost_norm);
  res9[0] = r.primal;
  res9[1] = r.dual;
  res9[2] = r.gap;
  res9[3] = r.pobj;
  res9[4] = r.dobj;
  res9[5] = r.abs_p;
  res9[6] = r.abs_d;
  res9[7] = r.gap_v;
  res9[8] = r.compl_;
}

// Plain products for the test hook and the power iteration. With `scale` set, the input is
// x / ctl->pw_len (the power iteration's normalisation, src/la/scaling.cpp:180).
template <typename V>
__global__ __launch_bounds__(kThreads) void k_spmv(PlanView<V> P, const double* x, double* y,
                                                   const Ctl* scale, double* part) {
  __shared__ double lds[kTile];
  __shared__ double red[4];
  const int4 item = P.items[blockIdx.x];
  if (scale && scale->pw_zero) return;
  PlainGather gather{x, 0.0, scale ? scale->pw_len : 0.0};
  PlainEpi epi{y};
  Acc<1, 0> acc;
  acc.zero();
  run_item<V, 1>(P, item, gather, epi, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc.s[0];
}

template <typename V>
__global__ __launch_bounds__(kTail) void k_spmv_tail(PlanView<V> P, double* y, const Ctl* scale,
                                                     double* lr_sum) {
  __shared__ double lds[kChunk];
  __shared__ double red[16];
  if (scale && scale->pw_zero) return;
  PlainEpi epi{y};
  Acc<1, 0> acc;
  acc.zero();
  finish_long_rows<V, 1>(P, epi, acc, lds);
  block_reduce(acc, red);
  if (threadIdx.x == 0) lr_sum[0] = acc.s[0];
}

// Anchor products of the mixed-precision delta form: y = P x in FP64 (x = a restart point), run
// after a restart has moved the restart point (ctl->restart_now) or when forced at start.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_anchor(PlanView<V> P, const double* x, double* y,
                                                     const Ctl* ctl, int force) {
  __shared__ double lds[kTile];
  __shared__ double red[4];
  const int4 item = P.items[blockIdx.x];
  if (!force && !ctl->restart_now) return;
  PlainGather gather{x, 0.0, 0.0};
  PlainEpi epi{y};
  Acc<1, 0> acc;
  acc.zero();
  run_item<V, 1>(P, item, gather, epi, acc, lds, red);
}

template <typename V>
__global__ __launch_bounds__(kTail) void k_anchor_tail(PlanView<V> P, double* y, const Ctl* ctl,
                                                       int force) {
  __shared__ double lds[kChunk];
  if (!force && !ctl->restart_now) return;
  PlainEpi epi{y};
  Acc<1, 0> acc;
  acc.zero();
  finish_long_rows<V, 1>(P, epi, acc, lds);
}

// Power-iteration step end (src/la/scaling.cpp:185-187): next = ||v||^2, norm = next^(1/4);
// the next step divides by length = sqrt(next).
__global__ __launch_bounds__(kThreads) void k_pw_finish(Ctl* ctl, const double* p, int np,
                                                        const double* lr_sum) {
  __shared__ double red[4];
  if (ctl->pw_zero) return;
  constexpr bool kSum1[1] = {false};
  double v[1];
  block_sums<1>(p, np, lr_sum, 1, 1, kSum1, v, red);
  if (threadIdx.x != 0) return;
  const double next = v[0];
  ctl->pw_norm = sqrt(sqrt(next));
  ctl->pw_len = sqrt(next);
  if (next == 0.0) ctl->pw_zero = 1;
}

// ---- preconditioning on the device (src/la/scaling.cpp:14-159) ---------------------------
// The same passes as the host preconditioner, bit for bit: a Ruiz round takes each row's and
// column's largest |a| (order-free), forms r = 1/sqrt(max) (1 for an empty line) and scales
// every entry as (a * r_i) * c_j in both stored copies; the Pock-Chambolle round sums |a| per
// row in ascending column order and per column in ascending row order, one lane per line,
// which is the order in which the reference's column sweep adds them.

// The wave-level helpers and the flat per-coordinate kernels are shared with the batched engine
// (scaling_waves.hpp).

// One workgroup per plan item. mode 0: largest |a| per row of a stream item into stat, per
// segment into seg (a long row's segments are combined by k_sc_longmax); mode 1: scale the
// item's entries. For the CSR the line factor is the row's, for the CSC the column's.
__global__ __launch_bounds__(kThreads) void k_sc_items(const int4* items, const int32_t* ptr,
                                                       const int32_t* idx, double* val,
                                                       const int4* lr_info, int mode, double* stat,
                                                       double* seg, const double* line_pass,
                                                       const double* idx_pass, int line_is_row) {
  __shared__ double red[kThreads / 64];
  const int4 it = items[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (it.y >= 0) {
    // A stream item's rows, L lanes each as in the products (L the largest power of two with
    // rows * L <= kThreads), so that all of them are in flight together: one wave per row
    // walked the Zipf LP's items of ~290 three-entry rows one memory round trip per row. The
    // maxima are exact in any order and each entry's (a * r) * c is computed as before, so the
    // results are unchanged bit for bit.
    const int rows = it.y - it.x;
    int L = 64;
    while (L > 1 && rows * L > kThreads) L >>= 1;
    const int grp = threadIdx.x / L, gl = threadIdx.x & (L - 1), ngrp = kThreads / L;
    for (int rr = grp; rr < rows; rr += ngrp) {  // the L lanes of a group stay together
      const int32_t row = it.x + rr;
      const int32_t k0 = ptr[row], k1 = ptr[row + 1];
      if (mode == 0) {
        double r = 0.0;
#pragma unroll 4
        for (int32_t k = k0 + gl; k < k1; k += L) r = smax(r, fabs(val[k]));
        for (int off = L >> 1; off > 0; off >>= 1) r = smax(r, __shfl_xor(r, off));
        if (gl == 0) stat[row] = r;
      } else {
        const double lf = line_pass[row];
#pragma unroll 4
        for (int32_t k = k0 + gl; k < k1; k += L) {
          const double other = idx_pass[idx[k]];
          const double r = line_is_row ? lf : other, c = line_is_row ? other : lf;
          val[k] = val[k] * r * c;
        }
      }
    }
    return;
  }
  // A segment: the workgroup's four waves take quarters of its entries.
  const int32_t k0 = it.z, k1 = it.w;
  const int32_t quarter = (k1 - k0 + kThreads / 64 - 1) / (kThreads / 64);
  const int32_t q0 = k0 + wave * quarter, q1 = q0 + quarter < k1 ? q0 + quarter : k1;
  if (mode == 0) {
    const double r = q0 < q1 ? wave_absmax(val, q0, q1, lane) : 0.0;
    if (lane == 0) red[wave] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int w = 1; w < kThreads / 64; ++w) m = smax(m, red[w]);
      seg[it.x] = m;
    }
  } else if (q0 < q1) {
    const int32_t row = lr_info[-it.y - 1].x;
    wave_apply(idx, val, q0, q1, lane, line_pass[row], idx_pass, line_is_row);
  }
}

// A long row's largest |a| from its segments' maxima (lr_info: row, first segment, count).
__global__ __launch_bounds__(kThreads) void k_sc_longmax(const int4* lr_info, int32_t nlr,
                                                         const double* seg, double* stat) {
  const int32_t lr = static_cast<int32_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (lr >= nlr) return;
  const int4 info = lr_info[lr];
  double m = 0.0;
  for (int32_t q = 0; q < info.z; ++q) m = smax(m, seg[info.y + q]);
  stat[info.x] = m;
}

// The Pock-Chambolle 1-norm of every line, one wave per line in storage order (a long row's
// entries too: its sum is one sequential chain).
__global__ __launch_bounds__(kThreads) void k_sc_linesum(const int32_t* ptr, const double* val,
                                                         int32_t lines, double* stat) {
  __shared__ double stage[kThreads * kScU];
  const int lane = threadIdx.x & 63;
  const int64_t line = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (line >= lines) return;  // the whole wave leaves together
  const double r = wave_abssum_ordered(val, ptr[line], ptr[line + 1], lane,
                                       stage + (threadIdx.x >> 6) * 64 * kScU);
  if (lane == 0) stat[line] = r;
}

// Smallest and largest |a| of the scaled matrix: per-workgroup extremes into part[2 g], then
// one workgroup combines them into out[0..1] (minimum and maximum are exact in any order).
__global__ __launch_bounds__(kThreads) void k_sc_extremes(const double* val, int64_t nnz,
                                                          double* part) {
  __shared__ double lo[kThreads], hi[kThreads];
  double a = nnz > 0 ? fabs(val[0]) : 0.0, z = 0.0;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < nnz;
       k += static_cast<int64_t>(gridDim.x) * kThreads) {
    const double mag = fabs(val[k]);
    a = smin(a, mag);
    z = smax(z, mag);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = z;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      lo[threadIdx.x] = smin(lo[threadIdx.x], lo[threadIdx.x + w]);
      hi[threadIdx.x] = smax(hi[threadIdx.x], hi[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lo[0];
    part[2 * blockIdx.x + 1] = hi[0];
  }
}

__global__ __launch_bounds__(kThreads) void k_sc_extremes_final(const double* part, int blocks,
                                                                double* out) {
  if (threadIdx.x != 0) return;
  double a = part[0], z = part[1];
  for (int b = 1; b < blocks; ++b) {
    a = smin(a, part[2 * b]);
    z = smax(z, part[2 * b + 1]);
  }
  out[0] = a;
  out[1] = z;
}

__global__ __launch_bounds__(kThreads) void k_sc_narrow(const double* v, float* f, int64_t count) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < count;
       k += static_cast<int64_t>(gridDim.x) * kThreads)
    f[k] = static_cast<float>(v[k]);
}

// ---- host side --------------------------------------------------------------------------

struct DevPlan {
  int32_t rows = 0, nitems = 0, nlr = 0, nseg = 0;
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  float* val32 = nullptr;   // FP32 values (mixed precision only)
  double* val64 = nullptr;  // FP64 values (always: ticks and reports evaluate in FP64)
  int4* items = nullptr;
  int4* lr_info = nullptr;
  int4* lr_groups = nullptr;
  int32_t ngroups = 0;
  int4 group0 = make_int4(0, 0, 0, 0);  // host copy of lr_groups[0]
  double* seg_part = nullptr;
  double* lr_in = nullptr;  // staged epilogue inputs of the long rows (kStage per long row)
  double* raw64 = nullptr;  // A's own values in this layout, kept for pdhg_dev_precondition
  template <typename V>
  PlanView<V> view() const {
    PlanView<V> p;
    p.rowptr = rowptr;
    p.col = col;
    if constexpr (sizeof(V) == 8) p.val = val64;
    else p.val = val32;
    p.items = items;
    p.lr_info = lr_info;
    p.lr_groups = lr_groups;
    p.seg_part = seg_part;
    p.nitems = nitems;
    p.nlr = nlr;
    p.ngroups = ngroups;
    p.group0 = group0;
    p.has_group0 = ngroups > 0 ? 1 : 0;
    return p;
  }
  // The view of a product kernel and its tail (an attempt's, or the tick's) that pass the long
  // rows' epilogue inputs through lr_in. The pairs run one after another on one stream, each
  // pair's staging read before the next pair writes.
  template <typename V>
  PlanView<V> staged_view() const {
    PlanView<V> p = view<V>();
    p.lr_in = lr_in;
    return p;
  }
};

}  // namespace sankhya_hip

// ---- scratch of the entry point that allocates nothing -----------------------------------
// pdhg_dev_evaluate_noalloc keeps its per-workgroup partial sums and long-row segment sums in
// this array: one copy per GPU, held until the process ends, shared by every call on that GPU.
// Host code never uses the name (a host-side stand-in): it uses the device address that
// hipGetSymbolAddress returns with the device current, recorded per device below.
constexpr size_t kScratchDoubles = size_t{1} << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

// With `device` current: its copy of the scratch array (null when the lookup fails).
double* scratch_for(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch[device] = static_cast<double*>(p);
  }
  return g_scratch[device];
}
}  // namespace

using namespace sankhya_hip;

struct pdhg_dev {
  int device = 0;
  int precision = PDHG_PRECISION_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  int fault_k = 0;
  int32_t m = 0, n = 0;
  int64_t nnz = 0;
  DevPlan A, T;
  int arith = PDHG_PRECISION_FP64;  // the iterates' current precision (mixed: FP32 until promoted)
  float *xb32 = nullptr, *yb32 = nullptr;    // iterate pairs, FP32 phase
  double *xb64 = nullptr, *yb64 = nullptr;   // iterate pairs, FP64
  double *xsum = nullptr, *ysum = nullptr, *xrest = nullptr, *yrest = nullptr;
  double *bestx = nullptr, *besty = nullptr, *sx = nullptr, *sy = nullptr, *sx2 = nullptr;
  double *axr = nullptr, *atyr = nullptr;  // Ahat xrest, Ahat^T yrest (mixed-precision delta form)
  double* sact = nullptr;
  double *chat = nullptr, *lhat = nullptr, *uhat = nullptr, *rlhat = nullptr, *ruhat = nullptr;
  double *dc = nullptr, *dr = nullptr, *cost = nullptr, *cl = nullptr, *cu = nullptr,
         *rl = nullptr, *ru = nullptr;
  double *part_p = nullptr, *lr_p = nullptr, *part_d = nullptr;
  double* stepf = nullptr;  // the step rule's two factors of the running attempt
  double *part_er = nullptr, *part_ec = nullptr, *lr_sum = nullptr;
  Ctl* ctl = nullptr;
  double* sc_part = nullptr;  // preconditioner scratch: per-workgroup extremes and the result
  int ruiz = 0;               // Ruiz passes of the last device preconditioning
  Ctl* ctl_host = nullptr;
  pdhg_tick* ticks = nullptr;
  hipGraphExec_t exec = nullptr;
  int exec_arith = -1;  // precision the captured graph was built for
  int chunk = kDefaultChunk;
  // Matrix loads of the iteration kernels cached rather than non-temporal: the handle's device
  // arrays fit the MALL with room to spare (kMallResident), so the matrices stay there between
  // attempts.
  bool cached = false;
  bool uploaded = false, started = false;
  int64_t bytes = 0;
  // Device memory comes in blocks of at least kBlockBytes, carved into 256-byte aligned pieces
  // and freed together with the handle: a few hipMalloc and hipFree calls per solve instead of
  // one per array (about 45). The 200k x 500k LP's release went from 6-7 ms to 2 ms.
  std::vector<void*> allocs;
  char* block_next = nullptr;
  size_t block_left = 0;
};

namespace {

int fail(pdhg_dev* h, int code, const std::string& what) {
  if (h) h->err = what;
  return code;
}

#define HIP_TRY(h, expr)                                                                  \
  do {                                                                                    \
    const hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                                 \
      return fail(h, PDHG_E_LAUNCH, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool injected(pdhg_dev* h) {
  if (h->fault_k <= 0) return false;
  if (--h->fault_k == 0) {
    h->err = "injected fault";
    return true;
  }
  return false;
}

// Every device array carries kSpare zeroed elements past its end. A lane outside its item
// gathers index 0 (pair element `cur` for the iterate pairs) and discards it; when the gathered
// vector is empty (the transpose of an LP with no rows, or an LP with no columns) that index
// lies past the vector's data, and the spare elements keep the read inside the allocation.
constexpr size_t kSpare = 2;

template <typename T>
int dalloc(pdhg_dev* h, T** p, size_t count) {
  *p = nullptr;
  const size_t bytes = (count + kSpare) * sizeof(T);
  const size_t take = (bytes + 255) & ~size_t{255};
  if (take > h->block_left) {
    const size_t block = std::max(take, kBlockBytes);
    void* b = nullptr;
    if (hipMalloc(&b, block) != hipSuccess)
      return fail(h, PDHG_E_ALLOC, "hipMalloc of " + std::to_string(block) + " bytes failed");
    h->allocs.push_back(b);
    h->block_next = static_cast<char*>(b);
    h->block_left = block;
  }
  *p = reinterpret_cast<T*>(h->block_next);
  h->block_next += take;
  h->block_left -= take;
  h->bytes += static_cast<int64_t>(bytes);
  HIP_TRY(h, hipMemsetAsync(*p + count, 0, kSpare * sizeof(T), h->stream));
  return PDHG_OK;
}

template <typename T>
int upload(pdhg_dev* h, T** dst, const T* src, size_t count) {
  int rc = dalloc(h, dst, count);
  if (rc) return rc;
  if (count && src) HIP_TRY(h, hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return PDHG_OK;
}

int make_plan(pdhg_dev* h, DevPlan* P, int32_t rows, int32_t ncols, int64_t nnz,
              const int32_t* rowptr, const int32_t* col, const double* val) {
  std::vector<int4> logical;
  const PlanGeo geo = single_geo(ncols, h->precision == PDHG_PRECISION_MIXED ? 8 : 16);
  const HostPlan hp = build_plan(rows, ncols, rowptr, col, &logical, geo);
  if (!plan_ok(hp, logical, rows, rowptr, geo))
    return fail(h, PDHG_E_ARGUMENT, "internal: work plan does not cover the matrix");
  P->rows = rows;
  P->nitems = static_cast<int32_t>(hp.items.size());
  P->nlr = static_cast<int32_t>(hp.lr_nseg.size());
  P->nseg = hp.nseg;
  int rc;
  if ((rc = upload(h, &P->rowptr, rowptr, static_cast<size_t>(rows) + 1))) return rc;
  // The kernels load nonzeros in aligned pairs, so the last pair of an odd-length array reads
  // one entry past nnz: the allocator's kSpare zeroed elements are that pad.
  static_assert(kSpare >= 2, "pair loads need a zeroed pad");
  const size_t un = static_cast<size_t>(nnz);
  if ((rc = upload(h, &P->col, col, un))) return rc;
  if (h->precision == PDHG_PRECISION_MIXED) {
    std::vector<float> f(un);
    for (int64_t k = 0; k < nnz; ++k) f[k] = static_cast<float>(val[k]);
    if ((rc = upload(h, &P->val32, f.data(), f.size()))) return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  if ((rc = upload(h, &P->val64, val, un))) return rc;
  if ((rc = upload(h, &P->items, hp.items.data(), hp.items.size()))) return rc;
  std::vector<int4> info(hp.lr_nseg.size());
  for (size_t k = 0; k < info.size(); ++k)
    info[k] = make_int4(hp.lr_row[k], hp.lr_base[k], hp.lr_nseg[k], 0);
  if ((rc = upload(h, &P->lr_info, info.data(), info.size()))) return rc;
  if ((rc = upload(h, &P->lr_groups, hp.groups.data(), hp.groups.size()))) return rc;
  P->ngroups = static_cast<int32_t>(hp.groups.size());
  if (!hp.groups.empty()) P->group0 = hp.groups[0];
  if ((rc = dalloc(h, &P->seg_part, static_cast<size_t>(hp.nseg) * 2))) return rc;
  if ((rc = dalloc(h, &P->lr_in, static_cast<size_t>(P->nlr) * kStage))) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

template <typename X>
Iter<X> iter_of(const pdhg_dev* h) {
  Iter<X> it;
  if constexpr (sizeof(X) == 8) {
    it.xbuf = h->xb64;
    it.ybuf = h->yb64;
  } else {
    it.xbuf = h->xb32;
    it.ybuf = h->yb32;
  }
  it.xsum = h->xsum;
  it.ysum = h->ysum;
  it.xrest = h->xrest;
  it.yrest = h->yrest;
  it.bestx = h->bestx;
  it.besty = h->besty;
  it.chat = h->chat;
  it.lhat = h->lhat;
  it.uhat = h->uhat;
  it.rlhat = h->rlhat;
  it.ruhat = h->ruhat;
  it.dc = h->dc;
  it.dr = h->dr;
  it.cost = h->cost;
  it.cl = h->cl;
  it.cu = h->cu;
  it.rl = h->rl;
  it.ru = h->ru;
  it.axr = h->axr;
  it.atyr = h->atyr;
  it.n = h->n;
  it.m = h->m;
  return it;
}

// What an attempt reads: both matrices in the iteration precision (row pointers, indices,
// values), the iterate pairs, and per column the cost, bounds and running sum, per row the
// bounds and running sum (the restart point and anchors in the FP32 phase).
int64_t attempt_bytes(const pdhg_dev* h) {
  const int64_t vb = h->precision == PDHG_PRECISION_MIXED ? 4 : 8;
  const int64_t m = h->m, n = h->n;
  return 2 * h->nnz * (4 + vb) + 4 * (m + n + 2) + n * (2 * vb + 4 * 8) + m * (2 * vb + 3 * 8) +
         (vb == 4 ? 16 * (m + n) : 0);
}
// The MI355X's MALL (Infinity Cache) holds 256 MB; an LP whose attempt reads at most this much
// runs its iteration kernels with cached matrix loads (the matrices stay resident between
// attempts), a larger one with non-temporal loads. The margin leaves room for the tick's and
// the report's arrays.
constexpr int64_t kMallResident = int64_t{200} << 20;

// Rounds of k_dual_tail's unconditional partial-sum loads (lane_sums_1_2_padded): the buffers
// part_p (one partial per item of Ahat^T) and part_d (two per item of Ahat) hold that many
// rounds of kSumU * kTail entries, zero past the items.
int sum_rounds(int64_t items_a, int64_t items_t) {
  const int64_t most = std::max<int64_t>(std::max(items_a, items_t), 1);
  constexpr int64_t kRound = int64_t{kSumU} * kTail;
  return static_cast<int>((most + kRound - 1) / kRound);
}

int apply_grid(const pdhg_dev* h) {
  const int64_t w = std::max<int64_t>(std::max<int64_t>(h->n, h->m), 1);
  return static_cast<int>(std::min<int64_t>((w + kThreads - 1) / kThreads, 8192));
}

// The two iteration kernels with the handle's matrix-load flavour (pdhg_dev::cached).
template <typename V, typename X>
void launch_primal(pdhg_dev* h, int force) {
  const Iter<X> it = iter_of<X>(h);
  if (h->cached)
    k_primal<V, X, false><<<h->T.nitems, kPrimalThreads, 0, h->stream>>>(h->T.staged_view<V>(), it,
                                                                   h->ctl, h->part_p, force);
  else
    k_primal<V, X, true><<<h->T.nitems, kPrimalThreads, 0, h->stream>>>(h->T.staged_view<V>(), it,
                                                                  h->ctl, h->part_p, force);
}
template <typename V, typename X>
void launch_dual(pdhg_dev* h, int force) {
  const Iter<X> it = iter_of<X>(h);
  if (h->cached)
    k_dual<V, X, false><<<h->A.nitems, kDualThreads, 0, h->stream>>>(h->A.staged_view<V>(), it, h->ctl,
                                                                 h->part_d, h->stepf, force);
  else
    k_dual<V, X, true><<<h->A.nitems, kDualThreads, 0, h->stream>>>(h->A.staged_view<V>(), it, h->ctl,
                                                                h->part_d, h->stepf, force);
}

template <typename V, typename X>
void launch_attempt(pdhg_dev* h, int force) {
  const Iter<X> it = iter_of<X>(h);
  if (h->T.nitems > 0) launch_primal<V, X>(h, force);
  if (h->T.nlr > 0)
    k_primal_tail<V, X><<<1, kTail, 0, h->stream>>>(h->T.staged_view<V>(), it, h->ctl, h->lr_p,
                                                    force);
  if (h->A.nitems > 0) launch_dual<V, X>(h, force);
  k_dual_tail<V, X><<<1, kTail, 0, h->stream>>>(h->A.staged_view<V>(), it, h->ctl, h->part_p,
                                                h->lr_p, h->T.nlr > 0 ? 1 : 0, h->part_d,
                                                h->A.nitems, sum_rounds(h->A.nitems, h->T.nitems),
                                                h->stepf, force);
}

// Anchor products Ahat xrest and Ahat^T yrest in FP64 (delta form), gated on a restart.
void launch_anchor(pdhg_dev* h, int force) {
  if (h->A.nitems > 0)
    k_anchor<double><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<double>(), h->xrest, h->axr,
                                                              h->ctl, force);
  if (h->A.nlr > 0)
    k_anchor_tail<double><<<1, kTail, 0, h->stream>>>(h->A.view<double>(), h->axr, h->ctl, force);
  if (h->T.nitems > 0)
    k_anchor<double><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<double>(), h->yrest,
                                                              h->atyr, h->ctl, force);
  if (h->T.nlr > 0)
    k_anchor_tail<double><<<1, kTail, 0, h->stream>>>(h->T.view<double>(), h->atyr, h->ctl, force);
}

template <typename V, typename X>
void launch_chunk(pdhg_dev* h) {
  const Iter<X> it = iter_of<X>(h);
  for (int s = 0; s < h->chunk; ++s) launch_attempt<V, X>(h, 0);
  // The evaluation tick reads the FP64 values in either precision: every convergence,
  // restart and primal-weight decision is taken on FP64 residuals.
  // Cached FP64 matrix loads when the FP64 matrices are the iteration's (the cached FP64 case).
  const bool cached64 = h->cached && h->precision == PDHG_PRECISION_FP64;
  if (h->A.nitems > 0) {
    if (cached64)
      k_eval_rows<double, X, false><<<h->A.nitems, kThreads, 0, h->stream>>>(
          h->A.staged_view<double>(), it, h->ctl, h->part_er);
    else
      k_eval_rows<double, X, true><<<h->A.nitems, kThreads, 0, h->stream>>>(
          h->A.staged_view<double>(), it, h->ctl, h->part_er);
  }
  if (h->T.nitems > 0) {
    if (cached64)
      k_eval_cols<double, X, false><<<h->T.nitems, kThreads, 0, h->stream>>>(
          h->T.staged_view<double>(), it, h->ctl, h->part_ec);
    else
      k_eval_cols<double, X, true><<<h->T.nitems, kThreads, 0, h->stream>>>(
          h->T.staged_view<double>(), it, h->ctl, h->part_ec);
  }
  k_decide<double, X><<<1, kTail, 0, h->stream>>>(h->A.staged_view<double>(),
                                                  h->T.staged_view<double>(), it, h->ctl,
                                                  h->part_er, h->part_ec,
                                                  sum_rounds(h->A.nitems, h->T.nitems), h->ticks);
  k_apply<X><<<apply_grid(h), kThreads, 0, h->stream>>>(it, h->ctl);
  if constexpr (kDelta<X>) launch_anchor(h, 0);
  k_clear<<<1, 64, 0, h->stream>>>(h->ctl);
}

// evaluate() of the point (bx, by) in original units: activity into act (may be null), reduced
// costs into red, the nine residuals into res9. The partial sums go to part_r (4 per item of
// Ahat) and part_c (4 per item of Ahat^T), the long-row segment sums to the views' seg_part.
template <typename V, typename X>
void launch_single_into(pdhg_dev* h, const PlanView<V>& A, const PlanView<V>& T, const double* bx,
                        const double* by, double* act, double* red, double* part_r, double* part_c,
                        double* res9) {
  const Iter<X> it = iter_of<X>(h);
  if (A.nitems > 0)
    k_single_rows<V, X><<<A.nitems, kThreads, 0, h->stream>>>(A, it, bx, by, act, part_r);
  if (T.nitems > 0)
    k_single_cols<V, X><<<T.nitems, kThreads, 0, h->stream>>>(T, it, bx, by, red, part_c);
  k_single_decide<V, X><<<1, kTail, 0, h->stream>>>(A, T, it, h->ctl, bx, by, act, red, part_r,
                                                    A.nitems, part_c, T.nitems, res9);
}

double* final_res_of(const pdhg_dev* h) {
  return reinterpret_cast<double*>(reinterpret_cast<char*>(h->ctl) + offsetof(Ctl, final_res));
}

template <typename V, typename X>
void launch_single(pdhg_dev* h, const double* bx, const double* by, double* act, double* red) {
  launch_single_into<V, X>(h, h->A.view<V>(), h->T.view<V>(), bx, by, act, red, h->part_er,
                           h->part_ec, final_res_of(h));
}

// y = P x through the same work items; long rows completed by the tail kernel. With `scale`
// the input is divided by the power iteration's current length.
template <typename V>
void launch_spmv(pdhg_dev* h, const DevPlan& P, const double* x, double* y, const Ctl* scale,
                 double* part, double* lr_sum) {
  if (P.nitems > 0)
    k_spmv<V><<<P.nitems, kThreads, 0, h->stream>>>(P.view<V>(), x, y, scale, part);
  k_spmv_tail<V><<<1, kTail, 0, h->stream>>>(P.view<V>(), y, scale, lr_sum);
}

// True while the iterates are FP32 (a mixed handle before pdhg_dev_promote).
bool mixed(const pdhg_dev* h) { return h->arith == PDHG_PRECISION_MIXED; }

// (Re)captures the graph chunk for the handle's current precision.
int capture(pdhg_dev* h) {
  if (h->exec) {
    HIP_TRY(h, hipGraphExecDestroy(h->exec));
    h->exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  if (mixed(h)) launch_chunk<float, float>(h);
  else launch_chunk<double, double>(h);
  HIP_TRY(h, hipStreamEndCapture(h->stream, &graph));
  const hipError_t e = hipGraphInstantiate(&h->exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(h, PDHG_E_LAUNCH, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->exec_arith = h->arith;
  return PDHG_OK;
}

int check_ready(pdhg_dev* h, bool need_start) {
  if (!h) return PDHG_E_ARGUMENT;
  if (!h->uploaded) return fail(h, PDHG_E_STATE, "no problem uploaded");
  if (need_start && !h->started) return fail(h, PDHG_E_STATE, "engine not started");
  if (injected(h)) return PDHG_E_INJECTED;
  return PDHG_OK;
}

// Ruiz rounds and one Pock-Chambolle round on the uploaded (unscaled) matrix, in place in
// both copies, then the scaled cost and bounds (and the FP32 copies in mixed precision).
// Scratch: sy and sact hold the row statistics and factors, sx and sx2 the column ones.
int device_scale(pdhg_dev* h, int ruiz_iterations, double* extremes) {
  const int32_t m = h->m, n = h->n;
  double *rstat = h->sy, *rpass = h->sact, *cstat = h->sx, *cpass = h->sx2;
  auto waves = [](int32_t lines) { return static_cast<unsigned>((lines + kThreads / 64 - 1) / (kThreads / 64)); };
  auto lanes = [](int64_t count) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + kThreads - 1) / kThreads, 8192)));
  };
  if (m > 0) k_sc_fill<<<lanes(m), kThreads, 0, h->stream>>>(h->dr, m, 1.0);
  if (n > 0) k_sc_fill<<<lanes(n), kThreads, 0, h->stream>>>(h->dc, n, 1.0);
  auto items = [&](const DevPlan& P, int mode, double* stat, const double* line_pass,
                   const double* idx_pass, int line_is_row) {
    if (P.nitems > 0)
      k_sc_items<<<P.nitems, kThreads, 0, h->stream>>>(P.items, P.rowptr, P.col, P.val64, P.lr_info,
                                                       mode, stat, P.seg_part, line_pass, idx_pass,
                                                       line_is_row);
    if (mode == 0 && P.nlr > 0)
      k_sc_longmax<<<(P.nlr + kThreads - 1) / kThreads, kThreads, 0, h->stream>>>(P.lr_info, P.nlr,
                                                                                  P.seg_part, stat);
  };
  auto factors_and_apply = [&]() {
    if (m > 0) k_sc_factor<<<lanes(m), kThreads, 0, h->stream>>>(rstat, rpass, h->dr, m);
    if (n > 0) k_sc_factor<<<lanes(n), kThreads, 0, h->stream>>>(cstat, cpass, h->dc, n);
    items(h->A, 1, nullptr, rpass, cpass, 1);
    items(h->T, 1, nullptr, cpass, rpass, 0);
  };
  for (int pass = 0; pass < ruiz_iterations; ++pass) {
    items(h->A, 0, rstat, nullptr, nullptr, 1);
    items(h->T, 0, cstat, nullptr, nullptr, 0);
    factors_and_apply();
  }
  if (m > 0) k_sc_linesum<<<waves(m), kThreads, 0, h->stream>>>(h->A.rowptr, h->A.val64, m, rstat);
  if (n > 0) k_sc_linesum<<<waves(n), kThreads, 0, h->stream>>>(h->T.rowptr, h->T.val64, n, cstat);
  factors_and_apply();
  const int64_t most = std::max<int64_t>(std::max<int64_t>(m, n), 1);
  k_sc_vectors<<<lanes(most), kThreads, 0, h->stream>>>(n, m, h->dc, h->dr, h->cost, h->cl, h->cu, h->rl,
                                                        h->ru, h->chat, h->lhat, h->uhat, h->rlhat,
                                                        h->ruhat);
  const int blocks = static_cast<int>(std::min<int64_t>(1024, std::max<int64_t>(1, (h->nnz + kThreads - 1) / kThreads)));
  if (!h->sc_part) {
    const int rc = dalloc(h, &h->sc_part, static_cast<size_t>(2 * 1024 + 2));
    if (rc) return rc;
  }
  double* part = h->sc_part;
  k_sc_extremes<<<blocks, kThreads, 0, h->stream>>>(h->A.val64, h->nnz, part);
  k_sc_extremes_final<<<1, kThreads, 0, h->stream>>>(part, blocks, part + 2 * blocks);
  if (h->precision == PDHG_PRECISION_MIXED && h->nnz > 0) {
    k_sc_narrow<<<lanes(h->nnz), kThreads, 0, h->stream>>>(h->A.val64, h->A.val32, h->nnz);
    k_sc_narrow<<<lanes(h->nnz), kThreads, 0, h->stream>>>(h->T.val64, h->T.val32, h->nnz);
  }
  HIP_TRY(h, hipGetLastError());
  double ext[2] = {0.0, 0.0};
  HIP_TRY(h, hipMemcpyAsync(ext, part + 2 * blocks, sizeof(ext), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (extremes) {
    extremes[0] = ext[0];
    extremes[1] = ext[1];
  }
  return PDHG_OK;
}

}  // namespace

extern "C" {

int pdhg_dev_probe(int device, char* desc, int desc_len) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) {
    if (desc && desc_len > 0) std::snprintf(desc, desc_len, "no HIP device %d", device);
    return PDHG_E_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    if (desc && desc_len > 0) std::snprintf(desc, desc_len, "device %d: properties unavailable", device);
    return PDHG_E_NO_DEVICE;
  }
  if (desc && desc_len > 0)
    std::snprintf(desc, desc_len, "%s (%s, %d CUs, %.1f GiB)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, static_cast<double>(prop.totalGlobalMem) / (1 << 30));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PDHG_E_NO_DEVICE;
  return PDHG_OK;
}

int pdhg_dev_create(int device, int precision, pdhg_dev** out) {
  if (!out) return PDHG_E_ARGUMENT;
  *out = nullptr;
  if (precision != PDHG_PRECISION_FP64 && precision != PDHG_PRECISION_MIXED) return PDHG_E_ARGUMENT;
  if (pdhg_dev_probe(device, nullptr, 0) != PDHG_OK) return PDHG_E_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PDHG_E_NO_DEVICE;
  pdhg_dev* h = new pdhg_dev();
  h->device = device;
  h->precision = precision;
  h->arith = precision;
  h->err.reserve(512);  // error texts of the no-allocation entry point fit without allocating
  (void)scratch_for(device);  // this device's scratch address, looked up once at set-up
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return PDHG_E_LAUNCH;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&h->ctl_host), sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
    hipStreamDestroy(h->stream);
    delete h;
    return PDHG_E_ALLOC;
  }
  std::memset(h->ctl_host, 0, sizeof(Ctl));
  *out = h;
  return PDHG_OK;
}

void pdhg_dev_destroy(pdhg_dev* h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->exec) hipGraphExecDestroy(h->exec);
  for (void* p : h->allocs) hipFree(p);
  if (h->ctl_host) hipHostFree(h->ctl_host);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

const char* pdhg_dev_last_error(const pdhg_dev* h) { return h ? h->err.c_str() : "null handle"; }

int64_t pdhg_dev_bytes(const pdhg_dev* h) { return h ? h->bytes : 0; }

void pdhg_dev_inject_fault(pdhg_dev* h, int k) {
  if (h) h->fault_k = k > 0 ? k : 0;
}

int64_t pdhg_dev_estimate_bytes(int32_t m, int32_t n, int64_t nnz, const int32_t* a_rowptr,
                                const int32_t* a_colidx, const int32_t* at_rowptr,
                                const int32_t* at_colidx, int precision, int unscaled,
                                int keep_unscaled) {
  if (m < 0 || n < 0 || nnz < 0 || nnz > 0x7fffffffLL || !a_rowptr || !at_rowptr) return -1;
  if (nnz > 0 && (!a_colidx || !at_colidx)) return -1;
  if (precision != PDHG_PRECISION_FP64 && precision != PDHG_PRECISION_MIXED) return -1;
  if (a_rowptr[m] != nnz || at_rowptr[n] != nnz) return -1;
  // The same sizes dalloc gives every array: count + kSpare elements.
  auto bytes = [](int64_t count, int64_t elem) { return (count + static_cast<int64_t>(kSpare)) * elem; };
  const bool mx = precision == PDHG_PRECISION_MIXED;
  int64_t total = 0;
  int32_t items_a = 0, items_t = 0;
  for (int side = 0; side < 2; ++side) {
    const int32_t rows = side == 0 ? m : n, cols = side == 0 ? n : m;
    const int32_t* rp = side == 0 ? a_rowptr : at_rowptr;
    const int32_t* ci = side == 0 ? a_colidx : at_colidx;
    std::vector<int4> logical;
    const HostPlan hp = build_plan(rows, cols, rp, ci, &logical, single_geo(cols, mx ? 8 : 16));
    const int64_t ni = static_cast<int64_t>(hp.items.size()), nl = static_cast<int64_t>(hp.lr_nseg.size());
    total += bytes(rows + 1, 4) + bytes(nnz, 4) + bytes(nnz, 8) + (mx ? bytes(nnz, 4) : 0);
    total += bytes(ni, 16) + bytes(nl, 16) + bytes(static_cast<int64_t>(hp.groups.size()), 16);
    total += bytes(2 * static_cast<int64_t>(hp.nseg), 8) + bytes(nl * kStage, 8);
    if (unscaled && keep_unscaled) total += bytes(nnz, 8);  // raw64
    (side == 0 ? items_a : items_t) = static_cast<int32_t>(ni);
  }
  const int64_t xe = mx ? 4 : 8;
  total += bytes(2 * int64_t{n}, xe) + bytes(2 * int64_t{m}, xe);     // iterate pairs
  if (mx) total += bytes(2 * int64_t{n}, 8) + bytes(2 * int64_t{m}, 8);  // FP64 pairs (promote)
  total += 6 * bytes(n, 8) + 6 * bytes(m, 8);                         // sums, restart, best, scratch
  total += 7 * bytes(n, 8) + 5 * bytes(m, 8);                         // problem vectors
  const int64_t ia = std::max(items_a, 1), it = std::max(items_t, 1);
  const int64_t padded = int64_t{sum_rounds(items_a, items_t)} * kSumU * kTail;
  total += bytes(padded, 8) + bytes(1, 8) + bytes(2 * padded, 8) + bytes(1, 8) +
           bytes(9 * padded, 8) + bytes(9 * padded, 8);
  total += bytes(1, sizeof(Ctl)) + bytes(kTickCap, sizeof(pdhg_tick)) + bytes(2, 8);
  if (unscaled) total += bytes(2 * 1024 + 2, 8);  // preconditioner scratch
  return total;
}

int64_t pdhg_dev_vram_reserve(int64_t total_bytes) {
  const int64_t kMinReserve = int64_t{256} << 20;
  return std::max(kMinReserve, total_bytes / 10);
}

int pdhg_dev_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return PDHG_E_ARGUMENT;
  if (pdhg_dev_probe(device, nullptr, 0) != PDHG_OK) return PDHG_E_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PDHG_E_NO_DEVICE;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return PDHG_E_LAUNCH;
  *free_bytes = static_cast<int64_t>(f);
  *total_bytes = static_cast<int64_t>(t);
  return PDHG_OK;
}

int pdhg_dev_upload(pdhg_dev* h, int32_t m, int32_t n, int64_t nnz, const int32_t* a_rowptr,
                    const int32_t* a_colidx, const double* a_val, const int32_t* at_rowptr,
                    const int32_t* at_colidx, const double* at_val, const double* cost_hat,
                    const double* lower_hat, const double* upper_hat,
                    const double* row_lower_hat, const double* row_upper_hat,
                    const double* col_scale, const double* row_scale, const double* min_cost,
                    const double* col_lower, const double* col_upper, const double* row_lower,
                    const double* row_upper, double bound_norm, double cost_norm) {
  if (!h) return PDHG_E_ARGUMENT;
  if (h->uploaded) return fail(h, PDHG_E_STATE, "problem already uploaded");
  if (m < 0 || n < 0 || nnz < 0 || nnz > 0x7fffffffLL) return fail(h, PDHG_E_ARGUMENT, "bad dimensions");
  if (!a_rowptr || !at_rowptr || (nnz > 0 && (!a_colidx || !a_val || !at_colidx || !at_val)))
    return fail(h, PDHG_E_ARGUMENT, "null matrix pointer");
  if (a_rowptr[m] != nnz || at_rowptr[n] != nnz || a_rowptr[0] != 0 || at_rowptr[0] != 0)
    return fail(h, PDHG_E_ARGUMENT, "row pointers do not match nnz");
  for (int32_t i = 0; i < m; ++i)
    if (a_rowptr[i + 1] < a_rowptr[i]) return fail(h, PDHG_E_ARGUMENT, "row pointer not monotone");
  for (int32_t j = 0; j < n; ++j)
    if (at_rowptr[j + 1] < at_rowptr[j]) return fail(h, PDHG_E_ARGUMENT, "transpose pointer not monotone");
  // One unsigned comparison per index (a negative index wraps above any bound), folded without
  // a branch so that the loop vectorises: 10M indices in well under a millisecond.
  {
    const uint32_t un = static_cast<uint32_t>(n), um = static_cast<uint32_t>(m);
    uint32_t bad = 0;
    for (int64_t k = 0; k < nnz; ++k)
      bad |= static_cast<uint32_t>(static_cast<uint32_t>(a_colidx[k]) >= un) |
             static_cast<uint32_t>(static_cast<uint32_t>(at_colidx[k]) >= um);
    if (bad) return fail(h, PDHG_E_ARGUMENT, "index out of range");
  }
  if (injected(h)) return PDHG_E_INJECTED;
  HIP_TRY(h, hipSetDevice(h->device));
  h->m = m;
  h->n = n;
  h->nnz = nnz;
  int rc;
  if ((rc = make_plan(h, &h->A, m, n, nnz, a_rowptr, a_colidx, a_val))) return rc;
  if ((rc = make_plan(h, &h->T, n, m, nnz, at_rowptr, at_colidx, at_val))) return rc;
  if (mixed(h)) {
    if ((rc = dalloc(h, &h->xb32, 2 * static_cast<size_t>(n))) ||
        (rc = dalloc(h, &h->yb32, 2 * static_cast<size_t>(m))))
      return rc;
  } else if ((rc = dalloc(h, &h->xb64, 2 * static_cast<size_t>(n))) ||
             (rc = dalloc(h, &h->yb64, 2 * static_cast<size_t>(m)))) {
    return rc;
  }
  const size_t un = static_cast<size_t>(n), um = static_cast<size_t>(m);
  double** nvec[] = {&h->xsum, &h->xrest, &h->bestx, &h->sx, &h->sx2, &h->atyr};
  for (double** v : nvec)
    if ((rc = dalloc(h, v, un))) return rc;
  double** mvec[] = {&h->ysum, &h->yrest, &h->besty, &h->sy, &h->sact, &h->axr};
  for (double** v : mvec)
    if ((rc = dalloc(h, v, um))) return rc;
  if ((rc = upload(h, &h->chat, cost_hat, un)) || (rc = upload(h, &h->lhat, lower_hat, un)) ||
      (rc = upload(h, &h->uhat, upper_hat, un)) || (rc = upload(h, &h->rlhat, row_lower_hat, um)) ||
      (rc = upload(h, &h->ruhat, row_upper_hat, um)) || (rc = upload(h, &h->dc, col_scale, un)) ||
      (rc = upload(h, &h->dr, row_scale, um)) || (rc = upload(h, &h->cost, min_cost, un)) ||
      (rc = upload(h, &h->cl, col_lower, un)) || (rc = upload(h, &h->cu, col_upper, un)) ||
      (rc = upload(h, &h->rl, row_lower, um)) || (rc = upload(h, &h->ru, row_upper, um)))
    return rc;
  const size_t ia = static_cast<size_t>(std::max(h->A.nitems, 1));
  const size_t it = static_cast<size_t>(std::max(h->T.nitems, 1));
  const size_t padded = static_cast<size_t>(sum_rounds(h->A.nitems, h->T.nitems)) * kSumU * kTail;
  if ((rc = dalloc(h, &h->part_p, padded)) || (rc = dalloc(h, &h->lr_p, 1)) ||
      (rc = dalloc(h, &h->part_d, 2 * padded)) || (rc = dalloc(h, &h->lr_sum, 1)) ||
      (rc = dalloc(h, &h->part_er, 9 * padded)) || (rc = dalloc(h, &h->part_ec, 9 * padded)) ||
      (rc = dalloc(h, &h->ctl, 1)) || (rc = dalloc(h, &h->ticks, static_cast<size_t>(kTickCap))) ||
      (rc = dalloc(h, &h->stepf, 2)))
    return rc;
  // The padding past the items must read zero (k_dual_tail and k_decide add it).
  HIP_TRY(h, hipMemsetAsync(h->part_p, 0, padded * sizeof(double), h->stream));
  HIP_TRY(h, hipMemsetAsync(h->part_d, 0, 2 * padded * sizeof(double), h->stream));
  HIP_TRY(h, hipMemsetAsync(h->part_er, 0, 9 * padded * sizeof(double), h->stream));
  HIP_TRY(h, hipMemsetAsync(h->part_ec, 0, 9 * padded * sizeof(double), h->stream));
  Ctl c;
  std::memset(&c, 0, sizeof(c));
  c.bound_norm = bound_norm;
  c.cost_norm = cost_norm;
  c.halt = 1;
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->cached = attempt_bytes(h) <= kMallResident;
  h->uploaded = true;
  return PDHG_OK;
}

int pdhg_dev_upload_unscaled(pdhg_dev* h, int32_t m, int32_t n, int64_t nnz,
                             const int32_t* a_rowptr, const int32_t* a_colidx, const double* a_val,
                             const int32_t* at_rowptr, const int32_t* at_colidx,
                             const double* at_val, const double* min_cost, const double* col_lower,
                             const double* col_upper, const double* row_lower,
                             const double* row_upper, double bound_norm, double cost_norm,
                             int ruiz_iterations, int keep_unscaled, double* extremes) {
  if (!h) return PDHG_E_ARGUMENT;
  if (ruiz_iterations < 0) return fail(h, PDHG_E_ARGUMENT, "negative Ruiz iteration count");
  if (!min_cost && n > 0) return fail(h, PDHG_E_ARGUMENT, "null cost");
  // The unscaled vectors stand in for the scaled ones until device_scale overwrites them.
  int rc = pdhg_dev_upload(h, m, n, nnz, a_rowptr, a_colidx, a_val, at_rowptr, at_colidx, at_val,
                           min_cost, col_lower, col_upper, row_lower, row_upper, nullptr, nullptr,
                           min_cost, col_lower, col_upper, row_lower, row_upper, bound_norm,
                           cost_norm);
  if (rc) return rc;
  if (keep_unscaled) {
    const size_t nz = static_cast<size_t>(nnz);
    if ((rc = dalloc(h, &h->A.raw64, nz)) || (rc = dalloc(h, &h->T.raw64, nz))) return rc;
    if (nz) {
      HIP_TRY(h, hipMemcpyAsync(h->A.raw64, h->A.val64, sizeof(double) * nz, hipMemcpyDeviceToDevice, h->stream));
      HIP_TRY(h, hipMemcpyAsync(h->T.raw64, h->T.val64, sizeof(double) * nz, hipMemcpyDeviceToDevice, h->stream));
    }
  }
  h->ruiz = ruiz_iterations;
  return device_scale(h, ruiz_iterations, extremes);
}

int pdhg_dev_precondition(pdhg_dev* h, double* extremes) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  if (!h->A.raw64 || !h->T.raw64)
    return fail(h, PDHG_E_STATE, "the unscaled values were not kept (keep_unscaled = 0)");
  HIP_TRY(h, hipSetDevice(h->device));
  const size_t nz = static_cast<size_t>(h->nnz);
  if (nz) {
    HIP_TRY(h, hipMemcpyAsync(h->A.val64, h->A.raw64, sizeof(double) * nz, hipMemcpyDeviceToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(h->T.val64, h->T.raw64, sizeof(double) * nz, hipMemcpyDeviceToDevice, h->stream));
  }
  h->started = false;
  return device_scale(h, h->ruiz, extremes);
}

int pdhg_dev_scaling(pdhg_dev* h, double* row, double* column, double* csc_values, double* cost_hat,
                     double* col_lower_hat, double* col_upper_hat, double* row_lower_hat,
                     double* row_upper_hat) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  HIP_TRY(h, hipSetDevice(h->device));
  const size_t un = static_cast<size_t>(h->n), um = static_cast<size_t>(h->m);
  const size_t nz = static_cast<size_t>(h->nnz);
  struct Item {
    double* dst;
    const double* src;
    size_t count;
  } items[] = {{row, h->dr, um},          {column, h->dc, un},       {csc_values, h->T.val64, nz},
               {cost_hat, h->chat, un},   {col_lower_hat, h->lhat, un}, {col_upper_hat, h->uhat, un},
               {row_lower_hat, h->rlhat, um}, {row_upper_hat, h->ruhat, um}};
  for (const Item& it : items)
    if (it.dst && it.count)
      HIP_TRY(h, hipMemcpyAsync(it.dst, it.src, sizeof(double) * it.count, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

int pdhg_dev_spmv(pdhg_dev* h, int transpose, const double* x, double* y) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  if (!x || !y) return fail(h, PDHG_E_ARGUMENT, "null vector");
  const DevPlan& P = transpose ? h->T : h->A;
  const int32_t in = transpose ? h->m : h->n;
  const int32_t out = transpose ? h->n : h->m;
  double* din = transpose ? h->sy : h->sx;
  double* dout = transpose ? h->sx2 : h->sy;
  HIP_TRY(h, hipMemcpyAsync(din, x, sizeof(double) * in, hipMemcpyHostToDevice, h->stream));
  double* part = transpose ? h->part_ec : h->part_er;
  if (mixed(h)) launch_spmv<float>(h, P, din, dout, nullptr, part, h->lr_sum);
  else launch_spmv<double>(h, P, din, dout, nullptr, part, h->lr_sum);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(y, dout, sizeof(double) * out, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

int pdhg_dev_spectral_norm(pdhg_dev* h, const double* v0, int iterations, double* norm) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  if (!norm || (h->n > 0 && !v0)) return fail(h, PDHG_E_ARGUMENT, "null argument");
  if (h->m == 0 || h->n == 0 || h->nnz == 0) {
    *norm = 1.0;
    return PDHG_OK;
  }
  // Reference order for the first length (src/la/scaling.cpp:176-178).
  double len = 0.0;
  for (int32_t j = 0; j < h->n; ++j) len += v0[j] * v0[j];
  len = std::sqrt(len);
  if (len == 0.0) {
    *norm = 1.0;
    return PDHG_OK;
  }
  Ctl c = {};
  HIP_TRY(h, hipMemcpyAsync(&c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  c.pw_len = len;
  c.pw_norm = 0.0;
  c.pw_zero = 0;
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->sx, v0, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  for (int step = 0; step < iterations; ++step) {
    launch_spmv<double>(h, h->A, h->sx, h->sy, h->ctl, h->part_er, h->lr_sum);
    launch_spmv<double>(h, h->T, h->sy, h->sx, nullptr, h->part_ec, h->lr_sum);
    k_pw_finish<<<1, kThreads, 0, h->stream>>>(h->ctl, h->part_ec, h->T.nitems, h->lr_sum);
    HIP_TRY(h, hipGetLastError());
  }
  HIP_TRY(h, hipMemcpyAsync(&c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  // A zero vector mid-iteration returns 1 (scaling.cpp:179).
  if (c.pw_zero) {
    *norm = 1.0;
    return PDHG_OK;
  }
  *norm = c.pw_norm > 0.0 ? c.pw_norm * 1.01 : 1.0;
  return PDHG_OK;
}

int pdhg_dev_start(pdhg_dev* h, const pdhg_params* p) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  if (!p || p->iteration_limit < 0) return fail(h, PDHG_E_ARGUMENT, "bad parameters");
  HIP_TRY(h, hipSetDevice(h->device));
  h->arith = h->precision;  // a mixed handle starts (again) in its FP32 phase
  Ctl c;
  HIP_TRY(h, hipMemcpyAsync(&c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const double bn = c.bound_norm, cn = c.cost_norm;
  std::memset(&c, 0, sizeof(c));
  c.bound_norm = bn;
  c.cost_norm = cn;
  c.eta = p->spectral_norm > 0.0 ? 1.0 / p->spectral_norm : 1.0;
  c.omega = 1.0;
  c.eta_ceiling = 1.0e3 / std::max(p->spectral_norm, 1e-12);
  c.restart_kkt = std::numeric_limits<double>::infinity();
  c.tolerance = p->tolerance;
  c.best_worst = std::numeric_limits<double>::infinity();
  c.best_primal = c.best_dual = std::numeric_limits<double>::infinity();
  c.iteration_limit = p->iteration_limit;
  c.use_restarts = p->restart ? 1 : 0;
  c.stop_at_request = p->stop_at_request ? 1 : 0;
  c.stop_rule = p->stop_rule == 1 ? 1 : 0;
  if (p->iteration_limit == 0) {
    c.halt = 1;
    c.limit_hit = 1;
  }
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  const int grid = apply_grid(h);
  if (mixed(h)) {
    k_init<float><<<grid, kThreads, 0, h->stream>>>(iter_of<float>(h));
    launch_anchor(h, 1);
  } else {
    k_init<double><<<grid, kThreads, 0, h->stream>>>(iter_of<double>(h));
  }
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int chunk = p->chunk_attempts > 0 ? p->chunk_attempts : kDefaultChunk;
  if (!h->exec || chunk != h->chunk || h->exec_arith != h->arith) {
    h->chunk = chunk;
    if ((rc = capture(h))) return rc;
  }
  h->started = true;
  return PDHG_OK;
}

int pdhg_dev_promote(pdhg_dev* h, const pdhg_params* p) {
  int rc = check_ready(h, true);
  if (rc) return rc;
  if (!p || p->iteration_limit < 0) return fail(h, PDHG_E_ARGUMENT, "bad parameters");
  if (h->precision != PDHG_PRECISION_MIXED || !mixed(h))
    return fail(h, PDHG_E_STATE, "promote needs a mixed-precision handle still in its FP32 phase");
  HIP_TRY(h, hipSetDevice(h->device));
  if (!h->xb64) {
    if ((rc = dalloc(h, &h->xb64, 2 * static_cast<size_t>(h->n))) ||
        (rc = dalloc(h, &h->yb64, 2 * static_cast<size_t>(h->m))))
      return rc;
  }
  const int64_t w = std::max<int64_t>(std::max<int64_t>(2 * int64_t{h->n}, 2 * int64_t{h->m}), 1);
  const int grid = static_cast<int>(std::min<int64_t>((w + kThreads - 1) / kThreads, 8192));
  k_widen<<<grid, kThreads, 0, h->stream>>>(h->xb32, h->xrest, h->xb64, 2 * int64_t{h->n},
                                            h->yb32, h->yrest, h->yb64, 2 * int64_t{h->m});
  HIP_TRY(h, hipGetLastError());
  Ctl c;
  HIP_TRY(h, hipMemcpyAsync(&c, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  c.tolerance = p->tolerance;
  c.stop_at_request = p->stop_at_request ? 1 : 0;
  c.stop_rule = p->stop_rule == 1 ? 1 : 0;
  c.iteration_limit = p->iteration_limit;
  c.done = 0;
  c.limit_hit = 0;
  c.halt = 0;
  if (c.iteration >= c.iteration_limit) {
    c.halt = 1;
    c.limit_hit = 1;
  }
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->arith = PDHG_PRECISION_FP64;
  return capture(h);
}

int pdhg_dev_run(pdhg_dev* h, int chunks, pdhg_state* st) {
  int rc = check_ready(h, true);
  if (rc) return rc;
  for (int c = 0; c < chunks; ++c) HIP_TRY(h, hipGraphLaunch(h->exec, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const Ctl& c = *h->ctl_host;
  if (st) {
    st->iteration = c.iteration;
    st->attempts = c.attempts;
    st->restarts = c.restarts;
    st->ticks = c.ntick;
    st->done = c.done;
    st->limit_hit = c.limit_hit;
    st->halted = c.halt;
    st->eta = c.eta;
    st->omega = c.omega;
    st->best_worst = c.best_worst;
    st->best_primal = c.best_primal;
    st->best_dual = c.best_dual;
    st->best_dual_objective = c.best_dobj;
  }
  return PDHG_OK;
}

int pdhg_dev_ticks(pdhg_dev* h, int64_t first, pdhg_tick* out, int64_t max, int64_t* count) {
  if (!h || !count) return PDHG_E_ARGUMENT;
  *count = 0;
  const int64_t total = h->ctl_host->ntick;
  int64_t begin = std::max<int64_t>(first, total - kTickCap);
  if (begin < 0) begin = 0;
  const int64_t want = std::min<int64_t>(total - begin, max);
  int64_t k = 0;
  while (k < want) {
    const int64_t slot = (begin + k) % kTickCap;
    const int64_t piece = std::min<int64_t>(want - k, kTickCap - slot);
    HIP_TRY(h, hipMemcpy(out + k, h->ticks + slot, sizeof(pdhg_tick) * piece, hipMemcpyDeviceToHost));
    k += piece;
  }
  *count = k;
  return PDHG_OK;
}

int pdhg_dev_finish(pdhg_dev* h, double* col_value, double* y, double* reduced, double* residuals) {
  int rc = check_ready(h, true);
  if (rc) return rc;
  launch_single<double, double>(h, h->bestx, h->besty, nullptr, h->sx2);
  HIP_TRY(h, hipGetLastError());
  if (col_value) HIP_TRY(h, hipMemcpyAsync(col_value, h->bestx, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  if (y) HIP_TRY(h, hipMemcpyAsync(y, h->besty, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  if (reduced) HIP_TRY(h, hipMemcpyAsync(reduced, h->sx2, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (residuals) std::memcpy(residuals, h->ctl_host->final_res, sizeof(double) * 9);
  return PDHG_OK;
}

int pdhg_dev_evaluate(pdhg_dev* h, const double* x, const double* y, double* activity,
                      double* reduced, double* residuals) {
  int rc = check_ready(h, false);
  if (rc) return rc;
  if ((h->n > 0 && !x) || (h->m > 0 && !y)) return fail(h, PDHG_E_ARGUMENT, "null vector");
  HIP_TRY(h, hipMemcpyAsync(h->sx, x, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->sy, y, sizeof(double) * h->m, hipMemcpyHostToDevice, h->stream));
  double* act = h->sact;
  launch_single<double, double>(h, h->sx, h->sy, act, h->sx2);
  HIP_TRY(h, hipGetLastError());
  if (activity) HIP_TRY(h, hipMemcpyAsync(activity, act, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  if (reduced) HIP_TRY(h, hipMemcpyAsync(reduced, h->sx2, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (residuals) std::memcpy(residuals, h->ctl_host->final_res, sizeof(double) * 9);
  return PDHG_OK;
}

int pdhg_dev_evaluate_noalloc(pdhg_dev* h, const double* x, const double* y, double* activity,
                              double* reduced, double* residuals) {
  // No allocation anywhere on this path, error texts included (h->err has room reserved).
  if (!h) return PDHG_E_ARGUMENT;
  auto refuse = [h](int code, const char* what) {
    h->err.assign(what);
    return code;
  };
  if (!h->uploaded) return refuse(PDHG_E_STATE, "pdhg_dev_evaluate_noalloc: no problem uploaded");
  if (injected(h)) return PDHG_E_INJECTED;
  if ((h->n > 0 && (!x || !reduced)) || (h->m > 0 && (!y || !activity)) || !residuals)
    return refuse(PDHG_E_ARGUMENT, "pdhg_dev_evaluate_noalloc: null device pointer");
  if (hipSetDevice(h->device) != hipSuccess)
    return refuse(PDHG_E_LAUNCH, "pdhg_dev_evaluate_noalloc: hipSetDevice failed");
  double* base = scratch_for(h->device);
  if (!base) return refuse(PDHG_E_LAUNCH, "pdhg_dev_evaluate_noalloc: hipGetSymbolAddress of the scratch array failed");
  const size_t na = static_cast<size_t>(h->A.nitems), nt = static_cast<size_t>(h->T.nitems);
  const size_t need = 4 * (na + nt) + static_cast<size_t>(h->A.nseg) + static_cast<size_t>(h->T.nseg);
  if (need > kScratchDoubles)
    return refuse(PDHG_E_ARGUMENT,
                  "pdhg_dev_evaluate_noalloc: the LP's work plans need more than the 1,048,576 "
                  "doubles (8 MiB) of scratch the library holds per GPU; use pdhg_dev_evaluate");
  // Layout: 4 partials per item of Ahat, 4 per item of Ahat^T, then the segment sums of each.
  PlanView<double> A = h->A.view<double>(), T = h->T.view<double>();
  A.seg_part = base + 4 * (na + nt);
  T.seg_part = A.seg_part + h->A.nseg;
  launch_single_into<double, double>(h, A, T, x, y, activity, reduced, base, base + 4 * na,
                                     residuals);
  if (hipGetLastError() != hipSuccess)
    return refuse(PDHG_E_LAUNCH, "pdhg_dev_evaluate_noalloc: kernel launch failed");
  if (hipStreamSynchronize(h->stream) != hipSuccess)
    return refuse(PDHG_E_LAUNCH, "pdhg_dev_evaluate_noalloc: the evaluation failed on the device");
  return PDHG_OK;
}

int pdhg_dev_time_kernels(pdhg_dev* h, int reps, double* primal_ms, double* dual_ms) {
  int rc = check_ready(h, true);
  if (rc) return rc;
  if (reps <= 0) return fail(h, PDHG_E_ARGUMENT, "reps must be positive");
  hipEvent_t e0, e1, e2;
  HIP_TRY(h, hipEventCreate(&e0));
  HIP_TRY(h, hipEventCreate(&e1));
  HIP_TRY(h, hipEventCreate(&e2));
  HIP_TRY(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps && h->T.nitems > 0; ++r) {
    if (mixed(h)) launch_primal<float, float>(h, 1);
    else launch_primal<double, double>(h, 1);
  }
  HIP_TRY(h, hipEventRecord(e1, h->stream));
  for (int r = 0; r < reps && h->A.nitems > 0; ++r) {
    if (mixed(h)) launch_dual<float, float>(h, 1);
    else launch_dual<double, double>(h, 1);
  }
  HIP_TRY(h, hipEventRecord(e2, h->stream));
  HIP_TRY(h, hipEventSynchronize(e2));
  float a = 0.f, b = 0.f;
  HIP_TRY(h, hipEventElapsedTime(&a, e0, e1));
  HIP_TRY(h, hipEventElapsedTime(&b, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  if (primal_ms) *primal_ms = a / reps;
  if (dual_ms) *dual_ms = b / reps;
  return PDHG_OK;
}

}  // extern "C"
