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
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "pdhg_kernels.hpp"

namespace sankhya_hip {

constexpr int64_t kEvalInterval = 40;  // kEvaluationInterval, src/pdhg/pdhg.cpp:65
constexpr int kDefaultChunk = 48;
constexpr int64_t kTickCap = 1 << 16;

// Device-resident control block. Written by the single-workgroup kernels, read by all.
struct Ctl {
  double eta, omega, eta_ceiling, restart_kkt, tolerance, bound_norm, cost_norm;
  double best_worst, best_primal, best_dual, best_dobj;
  double pw_len, pw_norm;
  double final_res[9];
  int64_t iteration, averaged, last_restart, restarts, iteration_limit, attempts, ntick;
  int64_t eval_count;
  int32_t cur, pending, tick_pending, halt, tick, apply, copy_best, chose_avg, restart_now;
  int32_t done, limit_hit, use_restarts, stop_at_request, no_info, pw_zero, pad;
};

template <typename X>
struct Iter {
  X* xbuf;  // 2n: (x_a, x_b) per column
  X* ybuf;  // 2m
  double *xsum, *ysum, *xrest, *yrest, *bestx, *besty;
  const double *chat, *lhat, *uhat, *rlhat, *ruhat, *dc, *dr, *cost, *cl, *cu, *rl, *ru;
  int32_t n, m;
};

template <typename X>
__device__ __forceinline__ void load_pair(const X* buf, int64_t j, int cur, double& a,
                                          double& b) {
  if constexpr (sizeof(X) == 8) {
    const double2 p = reinterpret_cast<const double2*>(buf)[j];
    a = cur ? p.y : p.x;
    b = cur ? p.x : p.y;
  } else {
    const float2 p = reinterpret_cast<const float2*>(buf)[j];
    a = static_cast<double>(cur ? p.y : p.x);
    b = static_cast<double>(cur ? p.x : p.y);
  }
}

// ---- iteration kernels ------------------------------------------------------------------

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_primal(PlanView<V> T, Iter<X> it, Ctl* ctl,
                                                     double* part, double* lrpart, int force) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 2];
  if (!force && ctl->halt) return;
  const int cur = ctl->cur, pend = ctl->pending;
  const double omega = ctl->omega;
  const double tau = ctl->eta / omega;
  auto gather = [&](int i) {
    RowSum<1> g;
    g.v[0] = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    return g;
  };
  auto epi = [&](int j, const RowSum<1>& s, Acc<1, 0>& a) {
    const double x = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    if (pend) it.xsum[j] += x;
    const double grad = it.chat[j] + s.v[0];
    const X stored = static_cast<X>(project(x - tau * grad, it.lhat[j], it.uhat[j]));
    it.xbuf[2 * (int64_t)j + (cur ^ 1)] = stored;
    const double d = static_cast<double>(stored) - x;
    a.s[0] += 0.5 * omega * d * d;
  };
  Acc<1, 0> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 1>(T, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc.s[0];
    if (lr >= 0) lrpart[lr] = lacc.s[0];
  }
}

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_dual(PlanView<V> A, Iter<X> it, Ctl* ctl,
                                                   double* part, double* lrpart, int force) {
  __shared__ double lds[2 * kTile];
  __shared__ double red[4 * 2];
  if (!force && ctl->halt) return;
  const int cur = ctl->cur, pend = ctl->pending;
  const double omega = ctl->omega;
  const double sigma = ctl->eta * omega;
  auto gather = [&](int j) {
    double x, xn;
    load_pair(it.xbuf, j, cur, x, xn);
    RowSum<2> g;
    g.v[0] = 2.0 * xn - x;
    g.v[1] = xn - x;
    return g;
  };
  auto epi = [&](int i, const RowSum<2>& s, Acc<2, 0>& a) {
    const double y = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    if (pend) it.ysum[i] += y;
    const double v = y + sigma * s.v[0];
    const X stored = static_cast<X>(v - sigma * project(v / sigma, it.rlhat[i], it.ruhat[i]));
    it.ybuf[2 * (int64_t)i + (cur ^ 1)] = stored;
    const double dy = static_cast<double>(stored) - y;
    a.s[0] += 0.5 * dy * dy / omega;
    a.s[1] += dy * s.v[1];
  };
  Acc<2, 0> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 2>(A, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc.s[0];
    part[2 * blockIdx.x + 1] = acc.s[1];
    if (lr >= 0) {
      lrpart[2 * lr] = lacc.s[0];
      lrpart[2 * lr + 1] = lacc.s[1];
    }
  }
}

// Adaptive step (src/pdhg/pdhg.cpp:293-335) and the tick test (:341-342).
__global__ __launch_bounds__(kThreads) void k_step(Ctl* ctl, const double* pp, int np,
                                                   const double* lrp, int nlrp,
                                                   const double* pd, int nd, const double* lrd,
                                                   int nlrd) {
  __shared__ double red[4];
  if (ctl->halt) return;
  const double mx = block_sum_strided(pp, np, 1, 0, red, false) +
                    block_sum_strided(lrp, nlrp, 1, 0, red, false);
  const double my = block_sum_strided(pd, nd, 2, 0, red, false) +
                    block_sum_strided(lrd, nlrd, 2, 0, red, false);
  const double inter = block_sum_strided(pd, nd, 2, 1, red, false) +
                       block_sum_strided(lrd, nlrd, 2, 1, red, false);
  if (threadIdx.x != 0) return;
  const double movement = mx + my;
  const double interaction = fabs(inter);
  const bool no_info = interaction <= 0.0;
  const double limit = no_info ? __builtin_huge_val() : movement / interaction;
  const int64_t iteration = ctl->iteration;
  const double exponent = static_cast<double>(iteration + 1 > 2 ? iteration + 1 : 2);
  const double shrink = 1.0 - pow(exponent, -0.3);
  const double grow = 1.0 + pow(exponent, -0.6);
  const double proposed = smin(shrink * limit, grow * ctl->eta);
  const double eta = ctl->eta;
  int64_t it = iteration;
  if (eta <= limit) {
    ctl->cur ^= 1;
    ++it;
    ctl->iteration = it;
    ctl->averaged += 1;
    ctl->pending = 1;
  } else {
    ctl->pending = 0;
  }
  if (!no_info) {
    double e = proposed;
    if (e < 1e-12) e = 1e-12;
    if (ctl->eta_ceiling < e) e = ctl->eta_ceiling;
    ctl->eta = e;
  }
  ctl->attempts += 1;
  ctl->no_info = no_info ? 1 : 0;
  const bool tick = it != 0 && (it % kEvalInterval == 0 || no_info);
  if (tick) {
    ctl->tick = 1;
    ctl->halt = 1;
  } else if (it >= ctl->iteration_limit) {
    ctl->halt = 1;
    ctl->limit_hit = 1;
  }
}

// ---- convergence test (src/pdhg/pdhg_evaluate.cpp:19-139), fused per row / column -------

__device__ __forceinline__ void row_terms(double act, double y, double rl, double ru,
                                          double& pv, double& sup, double& dv, double& cm) {
  double viol = 0.0;
  if (finite_bound(rl)) viol = smax(viol, rl - act);
  if (finite_bound(ru)) viol = smax(viol, act - ru);
  pv += viol * viol;
  if (y > 0.0) {
    if (finite_bound(ru)) sup += y * ru;
    else dv += y * y;
  } else if (y < 0.0) {
    if (finite_bound(rl)) sup += y * rl;
    else dv += y * y;
  }
  if (rl != ru) {
    const double ls = finite_bound(rl) ? act - rl : __builtin_huge_val();
    const double us = finite_bound(ru) ? ru - act : __builtin_huge_val();
    cm = smax(cm, fabs(y) * smin(ls, us));
  }
}

__device__ __forceinline__ void col_terms(double d, double x, double cl, double cu, double c,
                                          double& dv, double& bc, double& po, double& cm) {
  if (d > 0.0) {
    if (finite_bound(cl)) bc += d * cl;
    else dv += d * d;
  } else if (d < 0.0) {
    if (finite_bound(cu)) bc += d * cu;
    else dv += d * d;
  }
  if (cl != cu) {
    const double ls = finite_bound(cl) ? x - cl : __builtin_huge_val();
    const double us = finite_bound(cu) ? cu - x : __builtin_huge_val();
    cm = smax(cm, fabs(d) * smin(ls, us));
  }
  po += c * x;
}

// Tick: both candidates at once. Partials per workgroup (stride 9):
//   rows: pv_c sup_c dv_c pv_a sup_a dv_a dy_restart^2 | cm_c cm_a
//   cols: dv_c bc_c po_c dv_a bc_a po_a dx_restart^2   | cm_c cm_a
template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_eval_rows(PlanView<V> A, Iter<X> it, Ctl* ctl,
                                                        double* part, double* lrpart) {
  __shared__ double lds[2 * kTile];
  __shared__ double red[4 * 9];
  if (!ctl->tick) return;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = static_cast<double>(ctl->averaged > 0 ? ctl->averaged : 1);
  auto gather = [&](int j) {
    const double xc = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    RowSum<2> g;
    g.v[0] = xc;
    g.v[1] = (tp ? it.xsum[j] + xc : it.xsum[j]) / count;
    return g;
  };
  auto epi = [&](int i, const RowSum<2>& s, Acc<7, 2>& a) {
    const double dr = it.dr[i];
    const double yh = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    const double yah = (tp ? it.ysum[i] + yh : it.ysum[i]) / count;
    row_terms(s.v[0] / dr, yh * dr, it.rl[i], it.ru[i], a.s[0], a.s[1], a.s[2], a.m[0]);
    row_terms(s.v[1] / dr, yah * dr, it.rl[i], it.ru[i], a.s[3], a.s[4], a.s[5], a.m[1]);
    const double d = yh - it.yrest[i];
    a.s[6] += d * d;
  };
  Acc<7, 2> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 2>(A, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) part[9 * blockIdx.x + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) part[9 * blockIdx.x + 7 + k] = acc.m[k];
    if (lr >= 0) {
      for (int k = 0; k < 7; ++k) lrpart[9 * lr + k] = lacc.s[k];
      for (int k = 0; k < 2; ++k) lrpart[9 * lr + 7 + k] = lacc.m[k];
    }
  }
}

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_eval_cols(PlanView<V> T, Iter<X> it, Ctl* ctl,
                                                        double* part, double* lrpart) {
  __shared__ double lds[2 * kTile];
  __shared__ double red[4 * 9];
  if (!ctl->tick) return;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = static_cast<double>(ctl->averaged > 0 ? ctl->averaged : 1);
  auto gather = [&](int i) {
    const double yc = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    RowSum<2> g;
    g.v[0] = yc;
    g.v[1] = (tp ? it.ysum[i] + yc : it.ysum[i]) / count;
    return g;
  };
  auto epi = [&](int j, const RowSum<2>& s, Acc<7, 2>& a) {
    const double dc = it.dc[j];
    const double c = it.cost[j];
    const double xh = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    const double xah = (tp ? it.xsum[j] + xh : it.xsum[j]) / count;
    col_terms(c + s.v[0] / dc, xh * dc, it.cl[j], it.cu[j], c, a.s[0], a.s[1], a.s[2], a.m[0]);
    col_terms(c + s.v[1] / dc, xah * dc, it.cl[j], it.cu[j], c, a.s[3], a.s[4], a.s[5],
              a.m[1]);
    const double d = xh - it.xrest[j];
    a.s[6] += d * d;
  };
  Acc<7, 2> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 2>(T, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) part[9 * blockIdx.x + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) part[9 * blockIdx.x + 7 + k] = acc.m[k];
    if (lr >= 0) {
      for (int k = 0; k < 7; ++k) lrpart[9 * lr + k] = lacc.s[k];
      for (int k = 0; k < 2; ++k) lrpart[9 * lr + 7 + k] = lacc.m[k];
    }
  }
}

struct Res {
  double primal, dual, gap, pobj, dobj, abs_p, abs_d, gap_v, compl_;
};

__device__ inline Res make_res(double pv, double dv, double cm, double po, double dobj,
                               double bound_norm, double cost_norm) {
  Res r;
  r.abs_p = sqrt(pv);
  r.primal = r.abs_p / (1.0 + bound_norm);
  r.abs_d = sqrt(dv);
  r.dual = r.abs_d / (1.0 + cost_norm);
  r.compl_ = cm;
  r.pobj = po;
  r.dobj = dobj;
  const double ag = fabs(r.pobj - r.dobj);
  r.gap = ag / (1.0 + fabs(r.pobj) + fabs(r.dobj));
  r.gap_v = ag / smax(1.0, fabs(r.pobj));
  return r;
}
__device__ inline double worst(const Res& r) { return smax(smax(r.primal, r.dual), r.gap); }
__device__ inline bool meets_request(const Res& r, double t) {
  return r.primal <= t && r.dual <= t && r.gap <= t && r.abs_p <= 1e-7;
}
__device__ inline bool meets_standard(const Res& r) {
  return r.abs_p <= 1e-7 && r.abs_d <= 1e-7 && r.gap_v <= 1e-9 && r.compl_ <= 1e-6;
}

// Sums (idx < 7) or maxima (idx 7, 8) of a 9-wide partial array plus its long-row slots.
__device__ inline double part9(const double* p, int np, const double* lp, int nl, int idx,
                               double* red) {
  const bool mx = idx >= 7;
  const double a = block_sum_strided(p, np, 9, idx, red, mx);
  const double b = block_sum_strided(lp, nl, 9, idx, red, mx);
  return mx ? smax(a, b) : a + b;
}

// Best tracking, stopping test, restart and primal weight (src/pdhg/pdhg.cpp:344-483).
__global__ __launch_bounds__(kThreads) void k_decide(Ctl* ctl, const double* per, int nper,
                                                     const double* lper, int nlper,
                                                     const double* pec, int npec,
                                                     const double* lpec, int nlpec,
                                                     pdhg_tick* ticks) {
  __shared__ double red[4];
  __shared__ double v[18];
  if (!ctl->tick) return;
  for (int k = 0; k < 9; ++k) {
    const double a = part9(per, nper, lper, nlper, k, red);
    const double b = part9(pec, npec, lpec, nlpec, k, red);
    if (threadIdx.x == 0) {
      v[k] = a;
      v[9 + k] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  // rows: 0 pv_c 1 sup_c 2 dv_c 3 pv_a 4 sup_a 5 dv_a 6 dyr 7 cm_c 8 cm_a
  // cols: 9 dv_c 10 bc_c 11 po_c 12 dv_a 13 bc_a 14 po_a 15 dxr 16 cm_c 17 cm_a
  const double bn = ctl->bound_norm, cn = ctl->cost_norm;
  const Res cur = make_res(v[0], v[9] + v[2], smax(v[7], v[16]), v[11], v[10] - v[1], bn, cn);
  Res better = cur;
  int chose_avg = 0;
  if (ctl->averaged > 0) {
    const Res avg =
        make_res(v[3], v[12] + v[5], smax(v[8], v[17]), v[14], v[13] - v[4], bn, cn);
    if (worst(avg) < worst(cur)) {
      better = avg;
      chose_avg = 1;
    }
  }
  int copy = 0;
  if (worst(better) < ctl->best_worst) copy = 1;
  const bool stop = meets_request(better, ctl->tolerance) &&
                    (ctl->stop_at_request || meets_standard(better));
  int restarted = 0;
  if (stop) {
    copy = 1;
    ctl->done = 1;
  } else if (ctl->use_restarts) {
    const double kkt = worst(better);
    const int64_t since = ctl->iteration - ctl->last_restart;
    const bool sufficient = kkt <= 0.2 * ctl->restart_kkt;
    int64_t art = static_cast<int64_t>(0.36 * static_cast<double>(ctl->iteration));
    if (art < kEvalInterval) art = kEvalInterval;
    if (sufficient || since >= art) {
      const double dxn = sqrt(v[15]), dyn = sqrt(v[6]);
      if (dxn > 1e-12 && dyn > 1e-12) {
        double om = exp(0.5 * log(dyn / dxn) + (1.0 - 0.5) * log(ctl->omega));
        if (om < 1e-6) om = 1e-6;
        if (1e6 < om) om = 1e6;
        ctl->omega = om;
      }
      restarted = 1;
      ctl->restart_kkt = kkt;
      ctl->last_restart = ctl->iteration;
      ctl->restarts += 1;
    }
  }
  if (copy) {
    ctl->best_worst = worst(better);
    ctl->best_primal = better.primal;
    ctl->best_dual = better.dual;
    ctl->best_dobj = better.dobj;
  }
  pdhg_tick& t = ticks[ctl->ntick % kTickCap];
  t.iteration = ctl->iteration;
  t.primal_objective = better.pobj;
  t.primal = better.primal;
  t.dual = better.dual;
  t.gap = better.gap;
  t.restarted = restarted;
  t.chose_average = chose_avg;
  ctl->ntick += 1;
  ctl->eval_count = ctl->averaged;
  if (restarted) ctl->averaged = 0;
  ctl->tick_pending = ctl->pending;
  ctl->pending = 0;
  ctl->copy_best = copy;
  ctl->chose_avg = chose_avg;
  ctl->restart_now = restarted;
  ctl->apply = 1;
  ctl->tick = 0;
  if (ctl->done) {
    ctl->halt = 1;
  } else if (ctl->iteration >= ctl->iteration_limit) {
    ctl->halt = 1;
    ctl->limit_hit = 1;
  } else {
    ctl->halt = 0;
  }
}

template <typename X>
__global__ __launch_bounds__(kThreads) void k_apply(Iter<X> it, Ctl* ctl) {
  if (!ctl->apply) return;
  const int cur = ctl->cur, tp = ctl->tick_pending, cb = ctl->copy_best, avg = ctl->chose_avg,
            rs = ctl->restart_now;
  const double count = static_cast<double>(ctl->eval_count > 0 ? ctl->eval_count : 1);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t j = blockIdx.x * (int64_t)kThreads + threadIdx.x; j < it.n; j += stride) {
    const double xc = static_cast<double>(it.xbuf[2 * j + cur]);
    const double s = tp ? it.xsum[j] + xc : it.xsum[j];
    if (cb) it.bestx[j] = avg ? (s / count) * it.dc[j] : xc * it.dc[j];
    if (rs) {
      it.xsum[j] = 0.0;
      it.xrest[j] = xc;
    } else {
      it.xsum[j] = s;
    }
  }
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < it.m; i += stride) {
    const double yc = static_cast<double>(it.ybuf[2 * i + cur]);
    const double s = tp ? it.ysum[i] + yc : it.ysum[i];
    if (cb) it.besty[i] = avg ? (s / count) * it.dr[i] : yc * it.dr[i];
    if (rs) {
      it.ysum[i] = 0.0;
      it.yrest[i] = yc;
    } else {
      it.ysum[i] = s;
    }
  }
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
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < it.n; j += stride) {
    const double x0 = project(0.0, it.lhat[j], it.uhat[j]);
    it.xbuf[2 * j] = static_cast<X>(x0);
    it.xbuf[2 * j + 1] = static_cast<X>(0.0);
    it.xsum[j] = 0.0;
    it.xrest[j] = static_cast<double>(static_cast<X>(x0));
    // src/pdhg/pdhg.cpp:190: best_x starts as the scaled start point itself.
    it.bestx[j] = static_cast<double>(static_cast<X>(x0));
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < it.m; i += stride) {
    it.ybuf[2 * i] = static_cast<X>(0.0);
    it.ybuf[2 * i + 1] = static_cast<X>(0.0);
    it.ysum[i] = 0.0;
    it.yrest[i] = 0.0;
    it.besty[i] = 0.0;
  }
}

// Convergence test of one unscaled point (bx, by). Partials stride 4: 3 sums + 1 max.
template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_single_rows(PlanView<V> A, Iter<X> it,
                                                          const double* bx, const double* by,
                                                          double* act_out, double* part,
                                                          double* lrpart) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 4];
  auto gather = [&](int j) {
    RowSum<1> g;
    g.v[0] = bx[j] / it.dc[j];
    return g;
  };
  auto epi = [&](int i, const RowSum<1>& s, Acc<3, 1>& a) {
    const double act = s.v[0] / it.dr[i];
    if (act_out) act_out[i] = act;
    row_terms(act, by[i], it.rl[i], it.ru[i], a.s[0], a.s[1], a.s[2], a.m[0]);
  };
  Acc<3, 1> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 1>(A, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) part[4 * blockIdx.x + k] = acc.s[k];
    part[4 * blockIdx.x + 3] = acc.m[0];
    if (lr >= 0) {
      for (int k = 0; k < 3; ++k) lrpart[4 * lr + k] = lacc.s[k];
      lrpart[4 * lr + 3] = lacc.m[0];
    }
  }
}

template <typename V, typename X>
__global__ __launch_bounds__(kThreads) void k_single_cols(PlanView<V> T, Iter<X> it,
                                                          const double* bx, const double* by,
                                                          double* red_out, double* part,
                                                          double* lrpart) {
  __shared__ double lds[kTile];
  __shared__ double red[4 * 4];
  auto gather = [&](int i) {
    RowSum<1> g;
    g.v[0] = by[i] / it.dr[i];
    return g;
  };
  auto epi = [&](int j, const RowSum<1>& s, Acc<3, 1>& a) {
    const double d = it.cost[j] + s.v[0] / it.dc[j];
    if (red_out) red_out[j] = d;
    col_terms(d, bx[j], it.cl[j], it.cu[j], it.cost[j], a.s[0], a.s[1], a.s[2], a.m[0]);
  };
  Acc<3, 1> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 1>(T, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) part[4 * blockIdx.x + k] = acc.s[k];
    part[4 * blockIdx.x + 3] = acc.m[0];
    if (lr >= 0) {
      for (int k = 0; k < 3; ++k) lrpart[4 * lr + k] = lacc.s[k];
      lrpart[4 * lr + 3] = lacc.m[0];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_single_decide(Ctl* ctl, const double* pr, int npr,
                                                            const double* lpr, int nlpr,
                                                            const double* pc, int npc,
                                                            const double* lpc, int nlpc) {
  __shared__ double red[4];
  __shared__ double v[8];
  for (int k = 0; k < 4; ++k) {
    const bool mx = k == 3;
    const double a = block_sum_strided(pr, npr, 4, k, red, mx);
    const double b = block_sum_strided(lpr, nlpr, 4, k, red, mx);
    const double c = block_sum_strided(pc, npc, 4, k, red, mx);
    const double d = block_sum_strided(lpc, nlpc, 4, k, red, mx);
    if (threadIdx.x == 0) {
      v[k] = mx ? smax(a, b) : a + b;
      v[4 + k] = mx ? smax(c, d) : c + d;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  // rows: 0 pv 1 sup 2 dv 3 cm ; cols: 4 dv 5 bc 6 po 7 cm
  const Res r = make_res(v[0], v[4] + v[2], smax(v[3], v[7]), v[6], v[5] - v[1],
                         ctl->bound_norm, ctl->cost_norm);
  ctl->final_res[0] = r.primal;
  ctl->final_res[1] = r.dual;
  ctl->final_res[2] = r.gap;
  ctl->final_res[3] = r.pobj;
  ctl->final_res[4] = r.dobj;
  ctl->final_res[5] = r.abs_p;
  ctl->final_res[6] = r.abs_d;
  ctl->final_res[7] = r.gap_v;
  ctl->final_res[8] = r.compl_;
}

// Plain products for the test hook and the power iteration.
template <typename V>
__global__ __launch_bounds__(kThreads) void k_spmv(PlanView<V> P, const double* x, double* y,
                                                   const Ctl* scale_by_len, double* part,
                                                   double* lrpart) {
  __shared__ double lds[kTile];
  __shared__ double red[4];
  const double len = scale_by_len ? scale_by_len->pw_len : 1.0;
  if (scale_by_len && scale_by_len->pw_zero) return;
  auto gather = [&](int j) {
    RowSum<1> g;
    g.v[0] = scale_by_len ? x[j] / len : x[j];
    return g;
  };
  auto epi = [&](int i, const RowSum<1>& s, Acc<1, 0>& a) {
    y[i] = s.v[0];
    a.s[0] += s.v[0] * s.v[0];
  };
  Acc<1, 0> acc, lacc;
  acc.zero();
  lacc.zero();
  const int lr = run_item<V, 1>(P, gather, epi, acc, lacc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc.s[0];
    if (lr >= 0) lrpart[lr] = lacc.s[0];
  }
}

// Power-iteration step end (src/la/scaling.cpp:185-187): next = ||v||^2, norm = next^(1/4);
// the next step divides by length = sqrt(next).
__global__ __launch_bounds__(kThreads) void k_pw_finish(Ctl* ctl, const double* p, int np,
                                                        const double* lp, int nlp) {
  __shared__ double red[4];
  if (ctl->pw_zero) return;
  const double next = block_sum_strided(p, np, 1, 0, red, false) +
                      block_sum_strided(lp, nlp, 1, 0, red, false);
  if (threadIdx.x != 0) return;
  ctl->pw_norm = sqrt(sqrt(next));
  ctl->pw_len = sqrt(next);
  if (next == 0.0) ctl->pw_zero = 1;
}

// ---- host side --------------------------------------------------------------------------

struct DevPlan {
  int32_t rows = 0, nitems = 0, nlr = 0, nseg = 0;
  int32_t* rowptr = nullptr;
  int32_t* col = nullptr;
  void* val = nullptr;
  int4* items = nullptr;
  int32_t* lr_nseg = nullptr;
  int32_t* lr_base = nullptr;
  double* seg_part = nullptr;
  int32_t* lr_count = nullptr;
  template <typename V>
  PlanView<V> view() const {
    PlanView<V> p;
    p.rowptr = rowptr;
    p.col = col;
    p.val = static_cast<const V*>(val);
    p.items = items;
    p.lr_nseg = lr_nseg;
    p.lr_base = lr_base;
    p.seg_part = seg_part;
    p.lr_count = lr_count;
    p.nitems = nitems;
    p.nlr = nlr;
    return p;
  }
};

struct HostPlan {
  std::vector<int4> items;
  std::vector<int32_t> lr_nseg, lr_base;
  int32_t nseg = 0;
};

// Work items over a CSR row pointer: whole rows packed up to kTile nonzeros and kMaxRows
// rows per item; rows longer than kTile split into kSeg segments.
HostPlan build_plan(int32_t rows, const int32_t* rowptr) {
  HostPlan hp;
  int32_t r = 0;
  while (r < rows) {
    const int32_t len = rowptr[r + 1] - rowptr[r];
    if (len > kTile) {
      const int32_t lr = static_cast<int32_t>(hp.lr_nseg.size());
      const int32_t nseg = (len + kSeg - 1) / kSeg;
      hp.lr_nseg.push_back(nseg);
      hp.lr_base.push_back(hp.nseg);
      for (int32_t s = 0; s < nseg; ++s) {
        const int32_t k0 = rowptr[r] + s * kSeg;
        const int32_t k1 = std::min(rowptr[r] + (s + 1) * kSeg, rowptr[r + 1]);
        hp.items.push_back(make_int4(r, -lr - 1, k0, k1));
      }
      hp.nseg += nseg;
      ++r;
      continue;
    }
    const int32_t r0 = r;
    int64_t acc = 0;
    while (r < rows && r - r0 < kMaxRows) {
      const int32_t l = rowptr[r + 1] - rowptr[r];
      if (l > kTile || acc + l > kTile) break;
      acc += l;
      ++r;
    }
    hp.items.push_back(make_int4(r0, r, rowptr[r0], rowptr[r]));
  }
  return hp;
}

}  // namespace sankhya_hip

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
  void* xbuf = nullptr;
  void* ybuf = nullptr;
  double *xsum = nullptr, *ysum = nullptr, *xrest = nullptr, *yrest = nullptr;
  double *bestx = nullptr, *besty = nullptr, *sx = nullptr, *sy = nullptr, *sx2 = nullptr;
  double* sact = nullptr;
  double *chat = nullptr, *lhat = nullptr, *uhat = nullptr, *rlhat = nullptr, *ruhat = nullptr;
  double *dc = nullptr, *dr = nullptr, *cost = nullptr, *cl = nullptr, *cu = nullptr,
         *rl = nullptr, *ru = nullptr;
  double *part_p = nullptr, *lr_p = nullptr, *part_d = nullptr, *lr_d = nullptr;
  double *part_er = nullptr, *lr_er = nullptr, *part_ec = nullptr, *lr_ec = nullptr;
  Ctl* ctl = nullptr;
  Ctl* ctl_host = nullptr;
  pdhg_tick* ticks = nullptr;
  hipGraphExec_t exec = nullptr;
  int chunk = kDefaultChunk;
  bool uploaded = false, started = false;
  int64_t bytes = 0;
  std::vector<void*> allocs;
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

template <typename T>
int dalloc(pdhg_dev* h, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  const size_t bytes = count * sizeof(T);
  if (hipMalloc(reinterpret_cast<void**>(p), bytes) != hipSuccess) {
    *p = nullptr;
    return fail(h, PDHG_E_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  h->allocs.push_back(*p);
  h->bytes += static_cast<int64_t>(bytes);
  return PDHG_OK;
}

template <typename T>
int upload(pdhg_dev* h, T** dst, const T* src, size_t count) {
  int rc = dalloc(h, dst, count);
  if (rc) return rc;
  if (count && src) HIP_TRY(h, hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return PDHG_OK;
}

// Every row covered exactly once, stream items within the LDS tile and the row cap, long
// segments within kSeg: checked on the host before any kernel sees the plan.
bool plan_ok(const HostPlan& hp, int32_t rows, const int32_t* rowptr) {
  int32_t next_row = 0;
  for (const int4& it : hp.items) {
    if (it.y >= 0) {
      if (it.x != next_row || it.y <= it.x || it.y - it.x > kMaxRows) return false;
      if (it.z != rowptr[it.x] || it.w != rowptr[it.y] || it.w - it.z > kTile) return false;
      next_row = it.y;
    } else {
      const int32_t lr = -it.y - 1;
      if (lr < 0 || lr >= static_cast<int32_t>(hp.lr_nseg.size())) return false;
      if (it.z < rowptr[it.x] || it.w > rowptr[it.x + 1] || it.w <= it.z || it.w - it.z > kSeg) return false;
      if (it.w == rowptr[it.x + 1]) {
        if (it.x != next_row) return false;
        next_row = it.x + 1;
      }
    }
  }
  return next_row == rows;
}

int make_plan(pdhg_dev* h, DevPlan* P, int32_t rows, int64_t nnz, const int32_t* rowptr,
              const int32_t* col, const double* val) {
  const HostPlan hp = build_plan(rows, rowptr);
  if (!plan_ok(hp, rows, rowptr)) return fail(h, PDHG_E_ARGUMENT, "internal: work plan does not cover the matrix");
  P->rows = rows;
  P->nitems = static_cast<int32_t>(hp.items.size());
  P->nlr = static_cast<int32_t>(hp.lr_nseg.size());
  P->nseg = hp.nseg;
  int rc;
  if ((rc = upload(h, &P->rowptr, rowptr, static_cast<size_t>(rows) + 1))) return rc;
  if ((rc = upload(h, &P->col, col, static_cast<size_t>(nnz)))) return rc;
  if (h->precision == PDHG_PRECISION_MIXED) {
    std::vector<float> f(static_cast<size_t>(nnz));
    for (int64_t k = 0; k < nnz; ++k) f[k] = static_cast<float>(val[k]);
    float* d = nullptr;
    if ((rc = upload(h, &d, f.data(), f.size()))) return rc;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    P->val = d;
  } else {
    double* d = nullptr;
    if ((rc = upload(h, &d, val, static_cast<size_t>(nnz)))) return rc;
    P->val = d;
  }
  if ((rc = upload(h, &P->items, hp.items.data(), hp.items.size()))) return rc;
  if ((rc = upload(h, &P->lr_nseg, hp.lr_nseg.data(), hp.lr_nseg.size()))) return rc;
  if ((rc = upload(h, &P->lr_base, hp.lr_base.data(), hp.lr_base.size()))) return rc;
  if ((rc = dalloc(h, &P->seg_part, static_cast<size_t>(hp.nseg) * 2))) return rc;
  if ((rc = dalloc(h, &P->lr_count, static_cast<size_t>(P->nlr)))) return rc;
  HIP_TRY(h, hipMemsetAsync(P->lr_count, 0, sizeof(int32_t) * std::max<int32_t>(P->nlr, 1), h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

template <typename X>
Iter<X> iter_of(const pdhg_dev* h) {
  Iter<X> it;
  it.xbuf = static_cast<X*>(h->xbuf);
  it.ybuf = static_cast<X*>(h->ybuf);
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
  it.n = h->n;
  it.m = h->m;
  return it;
}

int apply_grid(const pdhg_dev* h) {
  const int64_t w = std::max<int64_t>(std::max<int64_t>(h->n, h->m), 1);
  return static_cast<int>(std::min<int64_t>((w + kThreads - 1) / kThreads, 8192));
}

template <typename V, typename X>
void launch_attempt(pdhg_dev* h, int force) {
  const Iter<X> it = iter_of<X>(h);
  if (h->T.nitems > 0)
    k_primal<V, X><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<V>(), it, h->ctl, h->part_p,
                                                            h->lr_p, force);
  if (h->A.nitems > 0)
    k_dual<V, X><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<V>(), it, h->ctl, h->part_d,
                                                          h->lr_d, force);
}

template <typename V, typename X>
void launch_chunk(pdhg_dev* h) {
  const Iter<X> it = iter_of<X>(h);
  for (int s = 0; s < h->chunk; ++s) {
    launch_attempt<V, X>(h, 0);
    k_step<<<1, kThreads, 0, h->stream>>>(h->ctl, h->part_p, h->T.nitems, h->lr_p, h->T.nlr,
                                          h->part_d, h->A.nitems, h->lr_d, h->A.nlr);
  }
  if (h->A.nitems > 0)
    k_eval_rows<V, X><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<V>(), it, h->ctl,
                                                               h->part_er, h->lr_er);
  if (h->T.nitems > 0)
    k_eval_cols<V, X><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<V>(), it, h->ctl,
                                                               h->part_ec, h->lr_ec);
  k_decide<<<1, kThreads, 0, h->stream>>>(h->ctl, h->part_er, h->A.nitems, h->lr_er, h->A.nlr,
                                          h->part_ec, h->T.nitems, h->lr_ec, h->T.nlr, h->ticks);
  k_apply<X><<<apply_grid(h), kThreads, 0, h->stream>>>(it, h->ctl);
  k_clear<<<1, 64, 0, h->stream>>>(h->ctl);
}

template <typename V, typename X>
void launch_single(pdhg_dev* h, const double* bx, const double* by, double* act, double* red) {
  const Iter<X> it = iter_of<X>(h);
  if (h->A.nitems > 0)
    k_single_rows<V, X><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<V>(), it, bx, by, act,
                                                                 h->part_er, h->lr_er);
  if (h->T.nitems > 0)
    k_single_cols<V, X><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<V>(), it, bx, by, red,
                                                                 h->part_ec, h->lr_ec);
  k_single_decide<<<1, kThreads, 0, h->stream>>>(h->ctl, h->part_er, h->A.nitems, h->lr_er,
                                                 h->A.nlr, h->part_ec, h->T.nitems, h->lr_ec,
                                                 h->T.nlr);
}

bool mixed(const pdhg_dev* h) { return h->precision == PDHG_PRECISION_MIXED; }

int check_ready(pdhg_dev* h, bool need_start) {
  if (!h) return PDHG_E_ARGUMENT;
  if (!h->uploaded) return fail(h, PDHG_E_STATE, "no problem uploaded");
  if (need_start && !h->started) return fail(h, PDHG_E_STATE, "engine not started");
  if (injected(h)) return PDHG_E_INJECTED;
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
  for (int64_t k = 0; k < nnz; ++k)
    if (a_colidx[k] < 0 || a_colidx[k] >= n || at_colidx[k] < 0 || at_colidx[k] >= m)
      return fail(h, PDHG_E_ARGUMENT, "index out of range");
  if (injected(h)) return PDHG_E_INJECTED;
  HIP_TRY(h, hipSetDevice(h->device));
  h->m = m;
  h->n = n;
  h->nnz = nnz;
  int rc;
  if ((rc = make_plan(h, &h->A, m, nnz, a_rowptr, a_colidx, a_val))) return rc;
  if ((rc = make_plan(h, &h->T, n, nnz, at_rowptr, at_colidx, at_val))) return rc;
  const size_t xs = mixed(h) ? sizeof(float) : sizeof(double);
  {
    char* p = nullptr;
    if ((rc = dalloc(h, &p, 2 * xs * std::max<int32_t>(n, 1)))) return rc;
    h->xbuf = p;
    if ((rc = dalloc(h, &p, 2 * xs * std::max<int32_t>(m, 1)))) return rc;
    h->ybuf = p;
  }
  const size_t un = static_cast<size_t>(n), um = static_cast<size_t>(m);
  double** nvec[] = {&h->xsum, &h->xrest, &h->bestx, &h->sx, &h->sx2};
  for (double** v : nvec)
    if ((rc = dalloc(h, v, un))) return rc;
  double** mvec[] = {&h->ysum, &h->yrest, &h->besty, &h->sy, &h->sact};
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
  const size_t la = static_cast<size_t>(std::max(h->A.nlr, 1));
  const size_t lt = static_cast<size_t>(std::max(h->T.nlr, 1));
  if ((rc = dalloc(h, &h->part_p, it)) || (rc = dalloc(h, &h->lr_p, lt)) ||
      (rc = dalloc(h, &h->part_d, 2 * ia)) || (rc = dalloc(h, &h->lr_d, 2 * la)) ||
      (rc = dalloc(h, &h->part_er, 9 * ia)) || (rc = dalloc(h, &h->lr_er, 9 * la)) ||
      (rc = dalloc(h, &h->part_ec, 9 * it)) || (rc = dalloc(h, &h->lr_ec, 9 * lt)) ||
      (rc = dalloc(h, &h->ctl, 1)) || (rc = dalloc(h, &h->ticks, static_cast<size_t>(kTickCap))))
    return rc;
  Ctl c;
  std::memset(&c, 0, sizeof(c));
  c.bound_norm = bound_norm;
  c.cost_norm = cost_norm;
  c.halt = 1;
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->uploaded = true;
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
  if (P.nitems > 0) {
    double* part = transpose ? h->part_ec : h->part_er;
    double* lrp = transpose ? h->lr_ec : h->lr_er;
    if (mixed(h))
      k_spmv<float><<<P.nitems, kThreads, 0, h->stream>>>(P.view<float>(), din, dout, nullptr, part, lrp);
    else
      k_spmv<double><<<P.nitems, kThreads, 0, h->stream>>>(P.view<double>(), din, dout, nullptr, part, lrp);
    HIP_TRY(h, hipGetLastError());
  }
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
    if (mixed(h)) {
      k_spmv<float><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<float>(), h->sx, h->sy, h->ctl, h->part_er, h->lr_er);
      k_spmv<float><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<float>(), h->sy, h->sx, nullptr, h->part_ec, h->lr_ec);
    } else {
      k_spmv<double><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<double>(), h->sx, h->sy, h->ctl, h->part_er, h->lr_er);
      k_spmv<double><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<double>(), h->sy, h->sx, nullptr, h->part_ec, h->lr_ec);
    }
    k_pw_finish<<<1, kThreads, 0, h->stream>>>(h->ctl, h->part_ec, h->T.nitems, h->lr_ec, h->T.nlr);
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
  if (p->iteration_limit == 0) {
    c.halt = 1;
    c.limit_hit = 1;
  }
  HIP_TRY(h, hipMemcpyAsync(h->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
  const int grid = apply_grid(h);
  if (mixed(h)) k_init<float><<<grid, kThreads, 0, h->stream>>>(iter_of<float>(h));
  else k_init<double><<<grid, kThreads, 0, h->stream>>>(iter_of<double>(h));
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int chunk = p->chunk_attempts > 0 ? p->chunk_attempts : kDefaultChunk;
  if (!h->exec || chunk != h->chunk) {
    if (h->exec) {
      hipGraphExecDestroy(h->exec);
      h->exec = nullptr;
    }
    h->chunk = chunk;
    hipGraph_t graph = nullptr;
    HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    if (mixed(h)) launch_chunk<float, float>(h);
    else launch_chunk<double, double>(h);
    HIP_TRY(h, hipStreamEndCapture(h->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&h->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(h, PDHG_E_LAUNCH, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  h->started = true;
  return PDHG_OK;
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
  if (mixed(h)) launch_single<float, float>(h, h->bestx, h->besty, nullptr, h->sx2);
  else launch_single<double, double>(h, h->bestx, h->besty, nullptr, h->sx2);
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
  if (mixed(h)) launch_single<float, float>(h, h->sx, h->sy, act, h->sx2);
  else launch_single<double, double>(h, h->sx, h->sy, act, h->sx2);
  HIP_TRY(h, hipGetLastError());
  if (activity) HIP_TRY(h, hipMemcpyAsync(activity, act, sizeof(double) * h->m, hipMemcpyDeviceToHost, h->stream));
  if (reduced) HIP_TRY(h, hipMemcpyAsync(reduced, h->sx2, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (residuals) std::memcpy(residuals, h->ctl_host->final_res, sizeof(double) * 9);
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
  const Iter<double> itd = iter_of<double>(h);
  const Iter<float> itf = iter_of<float>(h);
  HIP_TRY(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps && h->T.nitems > 0; ++r) {
    if (mixed(h)) k_primal<float, float><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<float>(), itf, h->ctl, h->part_p, h->lr_p, 1);
    else k_primal<double, double><<<h->T.nitems, kThreads, 0, h->stream>>>(h->T.view<double>(), itd, h->ctl, h->part_p, h->lr_p, 1);
  }
  HIP_TRY(h, hipEventRecord(e1, h->stream));
  for (int r = 0; r < reps && h->A.nitems > 0; ++r) {
    if (mixed(h)) k_dual<float, float><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<float>(), itf, h->ctl, h->part_d, h->lr_d, 1);
    else k_dual<double, double><<<h->A.nitems, kThreads, 0, h->stream>>>(h->A.view<double>(), itd, h->ctl, h->part_d, h->lr_d, 1);
  }
  HIP_TRY(h, hipEventRecord(e2, h->stream));
  HIP_TRY(h, hipEventSynchronize(e2));
  float a = 0.f, b = 0.f;
  HIP_TRY(h, hipEventElapsedTime(&a, e0, e1));
  HIP_TRY(h, hipEventElapsedTime(&b, e1, e2));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipEventDestroy(e2);
  if (primal_ms) *primal_ms = a / reps;
  if (dual_ms) *dual_ms = b / reps;
  return PDHG_OK;
}

}  // extern "C"
