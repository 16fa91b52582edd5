// pdhg_iterate.hpp - the restarted-PDHG iteration as device code, shared by the single-LP
// engine (pdhg_device.hip) and the batched engine (pdhg_batch.hip): the control block, the
// iterate view, the fused epilogues of the two products, the convergence-test terms, the
// adaptive step rule, and the tick decision (best point, stopping test, restart, primal weight).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pdhg_hip.h"
#include "pdhg_kernels.hpp"

namespace sankhya_hip {

constexpr int64_t kEvalInterval = 40;  // kEvaluationInterval, src/pdhg/pdhg.cpp:65
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
  int32_t done, limit_hit, use_restarts, stop_at_request, no_info, pw_zero, stop_rule;
};

// Iterate storage. FP64: xbuf/ybuf hold the iterates themselves. FP32 (mixed precision):
// they hold deltas from the restart point, x = xrest + dx, y = yrest + dy, and axr = Ahat xrest,
// atyr = Ahat^T yrest are kept in FP64 (recomputed after each restart). Rounding then scales
// with the distance from the restart point, which shrinks as the run converges, instead of
// with the iterate itself.
template <typename X>
constexpr bool kDelta = sizeof(X) == 4;

template <typename X>
struct Iter {
  X* xbuf;  // 2n: (x_a, x_b) per column
  X* ybuf;  // 2m
  double *xsum, *ysum, *xrest, *yrest, *bestx, *besty;
  const double *chat, *lhat, *uhat, *rlhat, *ruhat, *dc, *dr, *cost, *cl, *cu, *rl, *ru;
  const double *axr, *atyr;  // delta form only
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

// ---- epilogues and gathers, shared by the per-item kernels and the tail kernels ----------

template <typename X>
struct PrimalGather {  // y_cur[i]
  const X* ybuf;
  int cur;
  __device__ RowSum<1> operator()(int i) const {
    RowSum<1> g;
    g.v[0] = static_cast<double>(ybuf[2 * (int64_t)i + cur]);
    return g;
  }
};

// x' = proj(x - tau (c + Ahat^T y)); movement term (src/pdhg/pdhg.cpp:225-236,252-256).
template <typename X>
struct PrimalEpi {
  Iter<X> it;
  int cur, pend;
  double tau, omega;
  struct In {
    double x, xs, c, l, u, xr, g0;
  };
  __device__ In load(int j) const {
    In r;
    r.x = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    r.xs = pend ? it.xsum[j] : 0.0;
    r.c = it.chat[j];
    r.l = it.lhat[j];
    r.u = it.uhat[j];
    if constexpr (kDelta<X>) {
      r.xr = it.xrest[j];
      r.g0 = it.atyr[j];
    } else {
      r.xr = 0.0;
      r.g0 = 0.0;
    }
    return r;
  }
  __device__ void operator()(int j, const RowSum<1>& s, const In& in, Acc<1, 0>& a) const {
    if constexpr (kDelta<X>) {
      const double x = in.xr + in.x;
      if (pend) it.xsum[j] = in.xs + x;
      const double grad = in.c + (in.g0 + s.v[0]);
      const X stored = static_cast<X>(project(x - tau * grad, in.l, in.u) - in.xr);
      it.xbuf[2 * (int64_t)j + (cur ^ 1)] = stored;
      const double d = static_cast<double>(stored) - in.x;
      a.s[0] += 0.5 * omega * d * d;
    } else {
      if (pend) it.xsum[j] = in.xs + in.x;
      const double grad = in.c + s.v[0];
      const X stored = static_cast<X>(project(in.x - tau * grad, in.l, in.u));
      it.xbuf[2 * (int64_t)j + (cur ^ 1)] = stored;
      const double d = static_cast<double>(stored) - in.x;
      a.s[0] += 0.5 * omega * d * d;
    }
  }
};

template <typename X>
struct DualGather {  // (ext, dx) from one (x, x') pair
  const X* xbuf;
  int cur;
  __device__ RowSum<2> operator()(int j) const {
    double x, xn;
    load_pair(xbuf, j, cur, x, xn);
    RowSum<2> g;
    g.v[0] = 2.0 * xn - x;
    g.v[1] = xn - x;
    return g;
  }
};

// Moreau prox y' = v - sigma proj(v / sigma), movement and interaction (pdhg.cpp:238-277).
template <typename X>
struct DualEpi {
  Iter<X> it;
  int cur, pend;
  double sigma, omega;
  struct In {
    double y, ys, rl, ru, yr, g0;
  };
  __device__ In load(int i) const {
    In r;
    r.y = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    r.ys = pend ? it.ysum[i] : 0.0;
    r.rl = it.rlhat[i];
    r.ru = it.ruhat[i];
    if constexpr (kDelta<X>) {
      r.yr = it.yrest[i];
      r.g0 = it.axr[i];
    } else {
      r.yr = 0.0;
      r.g0 = 0.0;
    }
    return r;
  }
  __device__ void operator()(int i, const RowSum<2>& s, const In& in, Acc<2, 0>& a) const {
    if constexpr (kDelta<X>) {
      // ext = xrest + (2 dx' - dx), so Ahat ext = axr + s0; Ahat (x' - x) = s1 exactly.
      const double y = in.yr + in.y;
      if (pend) it.ysum[i] = in.ys + y;
      const double v = y + sigma * (in.g0 + s.v[0]);
      const X stored = static_cast<X>((v - sigma * project(v / sigma, in.rl, in.ru)) - in.yr);
      it.ybuf[2 * (int64_t)i + (cur ^ 1)] = stored;
      const double dy = static_cast<double>(stored) - in.y;
      a.s[0] += 0.5 * dy * dy / omega;
      a.s[1] += dy * s.v[1];
    } else {
      if (pend) it.ysum[i] = in.ys + in.y;
      const double v = in.y + sigma * s.v[0];
      const X stored = static_cast<X>(v - sigma * project(v / sigma, in.rl, in.ru));
      it.ybuf[2 * (int64_t)i + (cur ^ 1)] = stored;
      const double dy = static_cast<double>(stored) - in.y;
      a.s[0] += 0.5 * dy * dy / omega;
      a.s[1] += dy * s.v[1];
    }
  }
};

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

// Tick: both candidates (current, average) at once. Accumulators:
//   rows: s = pv_c sup_c dv_c pv_a sup_a dv_a dy_restart^2 | m = cm_c cm_a
//   cols: s = dv_c bc_c po_c dv_a bc_a po_a dx_restart^2   | m = cm_c cm_a
template <typename X>
struct TickRowsGather {  // x_cur[j] and the running average x_avg[j], scaled
  Iter<X> it;
  int cur, tp;
  double count;
  __device__ RowSum<2> operator()(int j) const {
    double xc = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    if constexpr (kDelta<X>) xc += it.xrest[j];
    RowSum<2> g;
    g.v[0] = xc;
    g.v[1] = (tp ? it.xsum[j] + xc : it.xsum[j]) / count;
    return g;
  }
};

template <typename X>
struct TickRowsEpi {
  Iter<X> it;
  int cur, tp;
  double count;
  struct In {
    double dr, yh, ys, rl, ru, yr;
  };
  __device__ In load(int i) const {
    In r;
    r.dr = it.dr[i];
    r.yh = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    r.ys = it.ysum[i];
    r.rl = it.rl[i];
    r.ru = it.ru[i];
    r.yr = it.yrest[i];
    if constexpr (kDelta<X>) r.yh += r.yr;
    return r;
  }
  __device__ void operator()(int i, const RowSum<2>& s, const In& in, Acc<7, 2>& a) const {
    const double yah = (tp ? in.ys + in.yh : in.ys) / count;
    row_terms(s.v[0] / in.dr, in.yh * in.dr, in.rl, in.ru, a.s[0], a.s[1], a.s[2], a.m[0]);
    row_terms(s.v[1] / in.dr, yah * in.dr, in.rl, in.ru, a.s[3], a.s[4], a.s[5], a.m[1]);
    const double d = in.yh - in.yr;
    a.s[6] += d * d;
  }
};

template <typename X>
struct TickColsGather {
  Iter<X> it;
  int cur, tp;
  double count;
  __device__ RowSum<2> operator()(int i) const {
    double yc = static_cast<double>(it.ybuf[2 * (int64_t)i + cur]);
    if constexpr (kDelta<X>) yc += it.yrest[i];
    RowSum<2> g;
    g.v[0] = yc;
    g.v[1] = (tp ? it.ysum[i] + yc : it.ysum[i]) / count;
    return g;
  }
};

template <typename X>
struct TickColsEpi {
  Iter<X> it;
  int cur, tp;
  double count;
  struct In {
    double dc, c, xh, xs, cl, cu, xr;
  };
  __device__ In load(int j) const {
    In r;
    r.dc = it.dc[j];
    r.c = it.cost[j];
    r.xh = static_cast<double>(it.xbuf[2 * (int64_t)j + cur]);
    r.xs = it.xsum[j];
    r.cl = it.cl[j];
    r.cu = it.cu[j];
    r.xr = it.xrest[j];
    if constexpr (kDelta<X>) r.xh += r.xr;
    return r;
  }
  __device__ void operator()(int j, const RowSum<2>& s, const In& in, Acc<7, 2>& a) const {
    const double xah = (tp ? in.xs + in.xh : in.xs) / count;
    col_terms(in.c + s.v[0] / in.dc, in.xh * in.dc, in.cl, in.cu, in.c, a.s[0], a.s[1], a.s[2],
              a.m[0]);
    col_terms(in.c + s.v[1] / in.dc, xah * in.dc, in.cl, in.cu, in.c, a.s[3], a.s[4], a.s[5],
              a.m[1]);
    const double d = in.xh - in.xr;
    a.s[6] += d * d;
  }
};

// One unscaled point (bx, by): rows s = pv sup dv | m = cm ; cols s = dv bc po | m = cm.
template <typename X>
struct SingleRowsGather {
  Iter<X> it;
  const double* bx;
  __device__ RowSum<1> operator()(int j) const {
    RowSum<1> g;
    g.v[0] = bx[j] / it.dc[j];
    return g;
  }
};
template <typename X>
struct SingleRowsEpi {
  Iter<X> it;
  const double* by;
  double* act_out;
  struct In {
    double dr, y, rl, ru;
  };
  __device__ In load(int i) const { return In{it.dr[i], by[i], it.rl[i], it.ru[i]}; }
  __device__ void operator()(int i, const RowSum<1>& s, const In& in, Acc<3, 1>& a) const {
    const double act = s.v[0] / in.dr;
    if (act_out) act_out[i] = act;
    row_terms(act, in.y, in.rl, in.ru, a.s[0], a.s[1], a.s[2], a.m[0]);
  }
};
template <typename X>
struct SingleColsGather {
  Iter<X> it;
  const double* by;
  __device__ RowSum<1> operator()(int i) const {
    RowSum<1> g;
    g.v[0] = by[i] / it.dr[i];
    return g;
  }
};
template <typename X>
struct SingleColsEpi {
  Iter<X> it;
  const double* bx;
  double* red_out;
  struct In {
    double c, dc, x, cl, cu;
  };
  __device__ In load(int j) const { return In{it.cost[j], it.dc[j], bx[j], it.cl[j], it.cu[j]}; }
  __device__ void operator()(int j, const RowSum<1>& s, const In& in, Acc<3, 1>& a) const {
    const double d = in.c + s.v[0] / in.dc;
    if (red_out) red_out[j] = d;
    col_terms(d, in.x, in.cl, in.cu, in.c, a.s[0], a.s[1], a.s[2], a.m[0]);
  }
};

// Plain product y = P x (optionally x / ||x|| for the power iteration), with sum of y^2.
struct PlainGather {
  const double* x;
  double inv;  // 1 / length, or 0 for none
  double len;
  __device__ RowSum<1> operator()(int j) const {
    RowSum<1> g;
    g.v[0] = len > 0.0 ? x[j] / len : x[j];
    return g;
  }
};
struct PlainEpi {
  double* y;
  struct In {};
  __device__ In load(int) const { return In{}; }
  __device__ void operator()(int i, const RowSum<1>& s, const In&, Acc<1, 0>& a) const {
    y[i] = s.v[0];
    a.s[0] += s.v[0] * s.v[0];
  }
};

template <typename X>
__device__ PrimalEpi<X> primal_epi(const Iter<X>& it, const Ctl* ctl) {
  return PrimalEpi<X>{it, ctl->cur, ctl->pending, ctl->eta / ctl->omega, ctl->omega};
}
template <typename X>
__device__ DualEpi<X> dual_epi(const Iter<X>& it, const Ctl* ctl) {
  return DualEpi<X>{it, ctl->cur, ctl->pending, ctl->eta * ctl->omega, ctl->omega};
}

// The control fields an attempt kernel reads, taken in one step at kernel start. The attempt
// kernels test `halt` before any work, and a workgroup must not wait for that test before
// issuing its other first loads (its work item, the kernel's own arguments): read_attempt keeps
// every one of these fields live across the test, so they are all loaded, and waited for, once.
struct AttemptCtl {
  double eta, omega;
  int32_t cur, pending, halt;
};
__device__ __forceinline__ AttemptCtl read_attempt(const Ctl* ctl) {
  return AttemptCtl{ctl->eta, ctl->omega, ctl->cur, ctl->pending, ctl->halt};
}
// True when the attempt kernels must exit at once. The other terms never hold (no work item
// has k1 < k0; omega and eta are positive, or NaN, which compares false; cur and pending are 0
// or 1): they make the test read the work item and every control field, so the compiler issues
// all those loads before the test instead of after it.
// `args` folds the kernel's argument pointers the same way (device allocations are 256-byte
// aligned, so their low bits are zero): their loads then go out with the first ones too.
__device__ __forceinline__ bool attempt_halted(const AttemptCtl& c, int force, const int4& item,
                                               uintptr_t args = 0) {
  // Bitwise, not short-circuit: every operand is evaluated, so no load waits behind a branch.
  const int never = static_cast<int>(item.w < item.z) | static_cast<int>(c.omega < 0.0) |
                    static_cast<int>(c.eta < -1.0) | static_cast<int>(c.cur < 0) |
                    static_cast<int>(c.pending < 0) | static_cast<int>((args & 7) != 0);
  const int halt = static_cast<int>(force == 0) & static_cast<int>(c.halt != 0);
  return (halt | never) != 0;
}
template <typename X>
__device__ __forceinline__ uintptr_t iter_bits(const Iter<X>& it) {
  auto b = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
  uintptr_t r = b(it.xbuf) | b(it.ybuf) | b(it.xsum) | b(it.ysum) | b(it.chat) | b(it.lhat) |
                b(it.uhat) | b(it.rlhat) | b(it.ruhat);
  if constexpr (kDelta<X>) r |= b(it.xrest) | b(it.yrest) | b(it.axr) | b(it.atyr);
  return r;
}
template <typename V>
__device__ __forceinline__ uintptr_t plan_bits(const PlanView<V>& P) {
  return reinterpret_cast<uintptr_t>(P.lr_info) | reinterpret_cast<uintptr_t>(P.seg_part) |
         reinterpret_cast<uintptr_t>(P.lr_in);
}
template <typename X>
__device__ PrimalEpi<X> primal_epi(const Iter<X>& it, const AttemptCtl& c) {
  return PrimalEpi<X>{it, c.cur, c.pending, c.eta / c.omega, c.omega};
}
template <typename X>
__device__ DualEpi<X> dual_epi(const Iter<X>& it, const AttemptCtl& c) {
  return DualEpi<X>{it, c.cur, c.pending, c.eta * c.omega, c.omega};
}
__device__ inline double tick_count(const Ctl* ctl) {
  return static_cast<double>(ctl->averaged > 0 ? ctl->averaged : 1);
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

// Adds a long-row accumulator into 9-wide totals (7 sums, 2 maxima).
__device__ inline void merge9(double (&v)[9], const Acc<7, 2>& a) {
  for (int k = 0; k < 7; ++k) v[k] += a.s[k];
  for (int k = 0; k < 2; ++k) v[7 + k] = smax(v[7 + k], a.m[k]);
}

// What the step rule reads from the control block, loaded by the deciding thread at kernel
// start so that read shares the kernel's first memory round trip.
struct StepSnap {
  int64_t iteration, iteration_limit, averaged, attempts;
  double eta, eta_ceiling;
  int32_t cur;
};
__device__ inline StepSnap step_snap(const Ctl* ctl) {
  return StepSnap{ctl->iteration, ctl->iteration_limit, ctl->averaged, ctl->attempts,
                  ctl->eta, ctl->eta_ceiling, ctl->cur};
}

// The step rule's factors 1 - k^-0.3 and 1 + k^-0.6 for k = max(2, iteration + 1)
// (src/pdhg/pdhg.cpp:303-306). They depend on the iteration count alone, so a kernel that runs
// before the decision can evaluate them (two double pow calls) off the decision's critical path.
struct StepFactors {
  double shrink, grow;
};
__device__ inline StepFactors step_factors(int64_t iteration) {
  const double exponent = static_cast<double>(iteration + 1 > 2 ? iteration + 1 : 2);
  return StepFactors{1.0 - pow(exponent, -0.3), 1.0 + pow(exponent, -0.6)};
}

// Adaptive step (src/pdhg/pdhg.cpp:293-335) and the tick test (:341-342), one thread, with the
// factors of step_factors(sn.iteration).
__device__ void step_decide(Ctl* ctl, const StepSnap& sn, double mx, double my, double inter,
                            const StepFactors& f) {
  const double movement = mx + my;
  const double interaction = fabs(inter);
  const bool no_info = interaction <= 0.0;
  const double limit = no_info ? __builtin_huge_val() : movement / interaction;
  const int64_t iteration = sn.iteration;
  const double shrink = f.shrink;
  const double grow = f.grow;
  const double proposed = smin(shrink * limit, grow * sn.eta);
  int64_t it = iteration;
  if (sn.eta <= limit) {
    ctl->cur = sn.cur ^ 1;
    ++it;
    ctl->iteration = it;
    ctl->averaged = sn.averaged + 1;
    ctl->pending = 1;
  } else {
    ctl->pending = 0;
  }
  if (!no_info) {
    double e = proposed;
    if (e < 1e-12) e = 1e-12;
    if (sn.eta_ceiling < e) e = sn.eta_ceiling;
    ctl->eta = e;
  }
  ctl->attempts = sn.attempts + 1;
  ctl->no_info = no_info ? 1 : 0;
  const bool tick = it != 0 && (it % kEvalInterval == 0 || no_info);
  if (tick) {
    ctl->tick = 1;
    ctl->halt = 1;
  } else if (it >= sn.iteration_limit) {
    ctl->halt = 1;
    ctl->limit_hit = 1;
  }
}

__device__ inline void step_decide(Ctl* ctl, const StepSnap& sn, double mx, double my, double inter) {
  step_decide(ctl, sn, mx, my, inter, step_factors(sn.iteration));
}

// Best tracking, stopping test, restart and primal weight (src/pdhg/pdhg.cpp:344-483) from a
// tick's 9-wide row and column totals, in one thread. Records the tick in `ticks` (a ring of
// kTickCap records) when that is not null.
__device__ inline void tick_decide(Ctl* ctl_mem, const double (&vr)[9], const double (&vc)[9],
                                   pdhg_tick* ticks) {
  // On a copy in registers: every field is loaded at once and stored once at the end. Worked
  // in place, each read of a field after a store to the block is its own memory round trip.
  Ctl local = *ctl_mem;
  Ctl* const ctl = &local;
  // rows: 0 pv_c 1 sup_c 2 dv_c 3 pv_a 4 sup_a 5 dv_a 6 dyr 7 cm_c 8 cm_a
  // cols: 0 dv_c 1 bc_c 2 po_c 3 dv_a 4 bc_a 5 po_a 6 dxr 7 cm_c 8 cm_a
  const double bn = ctl->bound_norm, cn = ctl->cost_norm;
  const Res rc = make_res(vr[0], vc[0] + vr[2], smax(vr[7], vc[7]), vc[2], vc[1] - vr[1], bn, cn);
  Res better = rc;
  int chose_avg = 0;
  if (ctl->averaged > 0) {
    const Res ravg = make_res(vr[3], vc[3] + vr[5], smax(vr[8], vc[8]), vc[5], vc[4] - vr[4], bn, cn);
    if (worst(ravg) < worst(rc)) {
      better = ravg;
      chose_avg = 1;
    }
  }
  int copy = 0;
  if (worst(better) < ctl->best_worst) copy = 1;
  // The reference's test (src/pdhg/pdhg.cpp:424-425), or, when asked for, the relative KKT
  // error alone (max of the relative primal, dual and gap residuals).
  const bool stop = ctl->stop_rule == 1
                        ? worst(better) <= ctl->tolerance
                        : meets_request(better, ctl->tolerance) &&
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
      const double dxn = sqrt(vc[6]), dyn = sqrt(vr[6]);
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
  if (ticks) {
    pdhg_tick& t = ticks[ctl->ntick % kTickCap];
    t.iteration = ctl->iteration;
    t.primal_objective = better.pobj;
    t.primal = better.primal;
    t.dual = better.dual;
    t.gap = better.gap;
    t.restarted = restarted;
    t.chose_average = chose_avg;
  }
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
  *ctl_mem = local;
}

// The tick's apply step for one column j / one row i (src/pdhg/pdhg.cpp:401-420,471-482): copy
// the chosen candidate into the best point (unscaled), and either fold the pending iterate into
// the running sum or, on a restart, clear the sum and move the restart point to the current
// iterate (the delta form also re-zeroes the current delta).
struct ApplyArgs {
  int cur, tp, cb, avg, rs;
  double count;
};
__device__ inline ApplyArgs apply_args(const Ctl* ctl) {
  return ApplyArgs{ctl->cur, ctl->tick_pending, ctl->copy_best, ctl->chose_avg, ctl->restart_now,
                   static_cast<double>(ctl->eval_count > 0 ? ctl->eval_count : 1)};
}
template <typename X>
__device__ inline void apply_col(const Iter<X>& it, const ApplyArgs& a, int64_t j) {
  double xc = static_cast<double>(it.xbuf[2 * j + a.cur]);
  if constexpr (kDelta<X>) xc += it.xrest[j];
  const double s = a.tp ? it.xsum[j] + xc : it.xsum[j];
  if (a.cb) it.bestx[j] = a.avg ? (s / a.count) * it.dc[j] : xc * it.dc[j];
  if (a.rs) {
    it.xsum[j] = 0.0;
    it.xrest[j] = xc;
    if constexpr (kDelta<X>) it.xbuf[2 * j + a.cur] = static_cast<X>(0.0);
  } else {
    it.xsum[j] = s;
  }
}
template <typename X>
__device__ inline void apply_row(const Iter<X>& it, const ApplyArgs& a, int64_t i) {
  double yc = static_cast<double>(it.ybuf[2 * i + a.cur]);
  if constexpr (kDelta<X>) yc += it.yrest[i];
  const double s = a.tp ? it.ysum[i] + yc : it.ysum[i];
  if (a.cb) it.besty[i] = a.avg ? (s / a.count) * it.dr[i] : yc * it.dr[i];
  if (a.rs) {
    it.ysum[i] = 0.0;
    it.yrest[i] = yc;
    if constexpr (kDelta<X>) it.ybuf[2 * i + a.cur] = static_cast<X>(0.0);
  } else {
    it.ysum[i] = s;
  }
}

// The start point (src/pdhg/pdhg.cpp:137-162,190): x = proj(0), y = 0, sums zero, the restart
// point and the best point at the start point. The delta form stores zero deltas around it.
template <typename X>
__device__ inline void init_col(const Iter<X>& it, int64_t j) {
  const double x0 = project(0.0, it.lhat[j], it.uhat[j]);
  it.xbuf[2 * j] = static_cast<X>(kDelta<X> ? 0.0 : x0);
  it.xbuf[2 * j + 1] = static_cast<X>(0.0);
  it.xsum[j] = 0.0;
  it.xrest[j] = x0;
  it.bestx[j] = x0;
}
template <typename X>
__device__ inline void init_row(const Iter<X>& it, int64_t i) {
  it.ybuf[2 * i] = static_cast<X>(0.0);
  it.ybuf[2 * i + 1] = static_cast<X>(0.0);
  it.ysum[i] = 0.0;
  it.yrest[i] = 0.0;
  it.besty[i] = 0.0;
}

}  // namespace sankhya_hip
