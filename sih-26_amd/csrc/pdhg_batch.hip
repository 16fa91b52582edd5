// pdhg_batch.hip - MI355X batched restarted-PDHG engine: many independent LPs, one workgroup
// per LP, each running whole chunks of attempts, evaluation ticks, restarts and its stopping
// test with workgroup barriers only. C ABI in include/pdhg_batch.h.
//
// Every LP follows the single-LP engine's algorithm (src/pdhg/pdhg.cpp:199-483), with the same
// device code for the epilogues, the step rule and the tick decision (pdhg_iterate.hpp) and the
// same tile walk for its products (pdhg_kernels.hpp), here at one of three workgroup
// geometries. What the batch buys is the schedule: no kernel boundary per attempt (a chunk of
// hundreds of attempts is one launch per geometry bin), and hundreds of LPs in flight at once
// instead of one small LP leaving most of the 256 CUs idle.
//
// Memory visibility inside a workgroup: epilogues write the next iterate to global memory and
// the next product gathers it after a workgroup barrier; all waves of a workgroup share one
// CU's L1, so the barrier orders them. The control block of the workgroup's LP lives in LDS for
// the whole launch (thread 0 writes it, every lane reads it after a barrier) and goes back to
// global memory at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pdhg_batch.h"
#include "pdhg_iterate.hpp"
#include "pdhg_kernels.hpp"
#include "batch_layout.hpp"
#include "plan_host.hpp"
#include "scaling_waves.hpp"

namespace sankhya_hip {
namespace batch {

// The three workgroup geometries a batch schedule picks from: threads, stream tile (8 nonzeros
// per thread), long-row segment, rows per stream item.
using Geo0 = Geometry<256, 2048, 2048, 512>;
using Geo1 = Geometry<512, 4096, 4096, 1024>;
using Geo2 = Geometry<1024, 8192, 8192, 2048>;
constexpr int kDefaultGeo = 2;
constexpr int kDefaultChunkAttempts = 480;
// Attempts per run launch before the tick kernel: one evaluation interval plus room for
// rejected attempts (the single-LP engine's chunk length).
constexpr int kSlotAttempts = 48;

static_assert(sizeof(Ctl) % 8 == 0, "control block is copied as 8-byte words");

template <typename V>
struct Mat {
  const int32_t* rowptr;
  const int32_t* col;
  const V* val;
};

struct PlanArrays {  // one matrix, one geometry
  const int4* items;
  const int4* lr_info;
  const int4* groups;
  double* seg_part;
};

template <typename V>
__device__ inline PlanView<V> lp_view(const Mat<V>& M, const PlanArrays& P, const LpPlan& p,
                                      int64_t nnz0, int32_t rp0) {
  PlanView<V> v;
  v.rowptr = M.rowptr + rp0;
  v.col = M.col + nnz0;
  v.val = M.val + nnz0;
  v.items = P.items + p.item0;
  v.lr_info = P.lr_info + p.lr0;
  v.lr_groups = P.groups + p.grp0;
  v.seg_part = P.seg_part + 2 * static_cast<int64_t>(p.seg0);
  v.nitems = p.nitems;
  v.nlr = p.nlr;
  v.ngroups = p.ngrp;
  v.has_group0 = 0;
  return v;
}

// One work item of a cooperative grid with what its workgroup needs of its LP: the item in its
// LP's numbering, its index among the geometry-0 items (its partial sums), the LP, and the LP's
// offsets into the concatenated arrays on the item's matrix side. Filled on the host when the
// grids are compacted, so that a workgroup's loads wait for this one uniform record instead of
// the item's index, then the item and its LP, then the LP's descriptor.
struct alignas(16) CoopItem {
  int4 item;
  int32_t b, lp, rp0, lr0;
  int64_t nnz0;
  int32_t seg0, row0, col0, m;
  int32_t n, pad;
};
static_assert(sizeof(CoopItem) == 64, "one 64-byte record per cooperative work item");

template <typename V>
__device__ inline PlanView<V> item_view(const Mat<V>& M, const PlanArrays& P, const CoopItem& r) {
  PlanView<V> v;
  v.rowptr = M.rowptr + r.rp0;
  v.col = M.col + r.nnz0;
  v.val = M.val + r.nnz0;
  // An item kernel is handed its item and walks no other: no item list or tail groups.
  v.items = P.items;
  v.lr_info = P.lr_info + r.lr0;
  v.lr_groups = P.groups;
  v.seg_part = P.seg_part + 2 * static_cast<int64_t>(r.seg0);
  v.nitems = 0;
  v.nlr = 0;
  v.ngroups = 0;
  v.has_group0 = 0;
  return v;
}

template <typename X>
__device__ inline Iter<X> lp_iter(const Iter<X>& g, int64_t r, int64_t c, int32_t m, int32_t n) {
  Iter<X> it = g;
  it.xbuf = g.xbuf + 2 * c;
  it.ybuf = g.ybuf + 2 * r;
  it.xsum = g.xsum + c;
  it.ysum = g.ysum + r;
  it.xrest = g.xrest + c;
  it.yrest = g.yrest + r;
  it.bestx = g.bestx + c;
  it.besty = g.besty + r;
  it.chat = g.chat + c;
  it.lhat = g.lhat + c;
  it.uhat = g.uhat + c;
  it.rlhat = g.rlhat + r;
  it.ruhat = g.ruhat + r;
  it.dc = g.dc + c;
  it.dr = g.dr + r;
  it.cost = g.cost + c;
  it.cl = g.cl + c;
  it.cu = g.cu + c;
  it.rl = g.rl + r;
  it.ru = g.ru + r;
  it.axr = g.axr + r;
  it.atyr = g.atyr + c;
  it.n = n;
  it.m = m;
  return it;
}
template <typename X>
__device__ inline Iter<X> lp_iter(const Iter<X>& g, const LpDesc& d) {
  return lp_iter(g, static_cast<int64_t>(d.row0), static_cast<int64_t>(d.col0), d.m, d.n);
}

// Everything a batch kernel reads, by value.
template <typename V, typename X>
struct Args {
  const LpDesc* desc;
  const int32_t* order;  // the bin's LPs, one per workgroup
  Ctl* ctl;              // nlp control blocks
  Mat<V> A, T;           // bulk values (FP32 in mixed precision)
  Mat<double> A64, T64;  // FP64 values: ticks, anchors, reports, power iteration
  PlanArrays pa, pt;     // this geometry's plans
  Iter<X> it;            // global iterate view (whole batch)
  double *axr, *atyr;    // writable anchor products (delta form)
  long long* clocks;     // per-LP device clock ticks spent in the last launch (tuning)
};

__device__ inline void ctl_load(Ctl* dst, const Ctl* src) {
  const int words = static_cast<int>(sizeof(Ctl) / 8);
  for (int k = threadIdx.x; k < words; k += blockDim.x)
    reinterpret_cast<int64_t*>(dst)[k] = reinterpret_cast<const int64_t*>(src)[k];
}
__device__ inline void ctl_store(Ctl* dst, const Ctl* src) {
  const int words = static_cast<int>(sizeof(Ctl) / 8);
  for (int k = threadIdx.x; k < words; k += blockDim.x)
    reinterpret_cast<int64_t*>(dst)[k] = reinterpret_cast<const int64_t*>(src)[k];
}

// All items of one plan, then its long rows, in this workgroup; barriers between items keep
// the LDS tile free for the next one.
template <class GEO, typename V, int NR, int NS, int NM, class G, class E>
__device__ inline void walk(const PlanView<V>& P, G& gather, E& epi, Acc<NS, NM>& acc,
                            double* lds, double* red) {
  for (int q = 0; q < P.nitems; ++q) {
    run_item_geo<GEO, V, NR>(P, P.items[q], gather, epi, acc, lds, red);
    __syncthreads();
  }
  if (P.nlr > 0) finish_long_rows_geo<GEO::kThreads, V, NR>(P, epi, acc, lds);
  __syncthreads();
}

template <class GEO>
constexpr int lds_doubles() {
  return GEO::kTile > 2 * kChunk ? GEO::kTile : 2 * kChunk;
}

// Anchor products of the delta form for one LP: axr = Ahat xrest, atyr = Ahat^T yrest.
template <class GEO, typename X>
__device__ inline void anchors(const PlanView<double>& A64, const PlanView<double>& T64,
                               const Iter<X>& it, double* axr, double* atyr, double* lds,
                               double* red) {
  Acc<1, 0> dummy;
  dummy.zero();
  PlainGather gx{it.xrest, 0.0, 0.0};
  PlainEpi ex{axr};
  walk<GEO, double, 1>(A64, gx, ex, dummy, lds, red);
  PlainGather gy{it.yrest, 0.0, 0.0};
  PlainEpi ey{atyr};
  walk<GEO, double, 1>(T64, gy, ey, dummy, lds, red);
}

// Up to `attempts` attempts of each LP in the bin (src/pdhg/pdhg.cpp:220-335); an LP stops
// early when the step rule raises a pending evaluation tick or hits the iteration limit.
template <class GEO, int G, typename V, typename X>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_run(Args<V, X> B, int attempts) {
  constexpr int NT = GEO::kThreads;
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(NT / 64) * 3];
  __shared__ Ctl sc;
  const long long t0 = wall_clock64();
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  ctl_load(&sc, B.ctl + lp);
  __syncthreads();
  if (sc.halt) return;
  const PlanView<V> A = lp_view(B.A, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<V> T = lp_view(B.T, B.pt, d.t[G], d.nnz0, d.trp0);
  const Iter<X> it = lp_iter(B.it, d);
  for (int a = 0; a < attempts && !sc.halt; ++a) {
    // Primal step over the rows of Ahat^T, then the dual step over the rows of Ahat.
    PrimalGather<X> gp{it.ybuf, sc.cur};
    PrimalEpi<X> ep = primal_epi(it, &sc);
    Acc<1, 0> ap;
    ap.zero();
    walk<GEO, V, 1>(T, gp, ep, ap, lds, red);
    DualGather<X> gd{it.xbuf, sc.cur};
    DualEpi<X> ed = dual_epi(it, &sc);
    Acc<3, 0> r;
    {
      Acc<2, 0> ad;
      ad.zero();
      walk<GEO, V, 2>(A, gd, ed, ad, lds, red);
      r.s[0] = ap.s[0];
      r.s[1] = ad.s[0];
      r.s[2] = ad.s[1];
    }
    block_reduce(r, red);
    if (threadIdx.x == 0) {
      const StepSnap sn = step_snap(&sc);
      step_decide(&sc, sn, r.s[0], r.s[1], r.s[2]);
    }
    __syncthreads();
  }
  ctl_store(B.ctl + lp, &sc);
  if (threadIdx.x == 0) B.clocks[lp] += wall_clock64() - t0;
}

// The evaluation tick of each LP that has one pending (src/pdhg/pdhg.cpp:341-483): both
// candidates on FP64 values, the decision, the apply step, and the delta form's new anchors
// after a restart.
template <class GEO, int G, typename X>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_tick(Args<double, X> B) {
  constexpr int NT = GEO::kThreads;
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(NT / 64) * 9];
  __shared__ double tot[18];
  __shared__ Ctl sc;
  const long long t0 = wall_clock64();
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  ctl_load(&sc, B.ctl + lp);
  __syncthreads();
  if (!sc.tick) return;
  const PlanView<double> A64 = lp_view(B.A64, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(B.T64, B.pt, d.t[G], d.nnz0, d.trp0);
  const Iter<X> it = lp_iter(B.it, d);
  const int cur = sc.cur, tp = sc.pending;
  const double count = tick_count(&sc);
  {
    TickRowsGather<X> gr{it, cur, tp, count};
    TickRowsEpi<X> er{it, cur, tp, count};
    Acc<7, 2> ra;
    ra.zero();
    walk<GEO, double, 2>(A64, gr, er, ra, lds, red);
    block_reduce(ra, red);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 7; ++k) tot[k] = ra.s[k];
      tot[7] = ra.m[0];
      tot[8] = ra.m[1];
    }
  }
  {
    TickColsGather<X> gc{it, cur, tp, count};
    TickColsEpi<X> ec{it, cur, tp, count};
    Acc<7, 2> ca;
    ca.zero();
    walk<GEO, double, 2>(T64, gc, ec, ca, lds, red);
    block_reduce(ca, red);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 7; ++k) tot[9 + k] = ca.s[k];
      tot[16] = ca.m[0];
      tot[17] = ca.m[1];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double vr[9], vc[9];
    for (int k = 0; k < 9; ++k) {
      vr[k] = tot[k];
      vc[k] = tot[9 + k];
    }
    tick_decide(&sc, vr, vc, nullptr);
  }
  __syncthreads();
  const ApplyArgs aa = apply_args(&sc);
  for (int64_t j = threadIdx.x; j < it.n; j += NT) apply_col(it, aa, j);
  for (int64_t i = threadIdx.x; i < it.m; i += NT) apply_row(it, aa, i);
  __syncthreads();
  if constexpr (kDelta<X>) {
    if (sc.restart_now) anchors<GEO>(A64, T64, it, B.axr + d.row0, B.atyr + d.col0, lds, red);
  }
  if (threadIdx.x == 0) {
    sc.apply = 0;
    sc.copy_best = 0;
    sc.restart_now = 0;
  }
  __syncthreads();
  ctl_store(B.ctl + lp, &sc);
  if (threadIdx.x == 0) B.clocks[lp] += wall_clock64() - t0;
}

template <typename X>
__global__ void k_batch_init(Iter<X> it) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < it.n; j += stride)
    init_col(it, j);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < it.m; i += stride)
    init_row(it, i);
}

// ---- preconditioning of every LP on the device (src/la/scaling.cpp:14-159) ---------------
// The single-LP engine's passes over the concatenated batch, one wave per line (row of A or
// column of A, i.e. row of A^T) of every LP: the line's LP comes from row_lp / col_lp, its
// entries from the LP's own pointers, and the other side's index is offset by the LP's first
// column (row). Results equal the host passes of each LP bit for bit (scaling_waves.hpp).
struct BatchLines {
  const int32_t* ptr;   // a_rowptr (side 0) or t_rowptr (side 1), per-LP local pointers
  const int32_t* idx;   // a_col / t_col, local indices of the other side
  double* val;          // a_val64 / t_val64
  const LpDesc* desc;
  const int32_t* line_lp;
  int64_t lines;
  int32_t side;
};

__device__ __forceinline__ void batch_line(const BatchLines& L, int64_t line, int32_t& k0,
                                           int32_t& k1, int64_t& base, int32_t& other0) {
  const LpDesc& d = L.desc[L.line_lp[line]];
  const int32_t local = static_cast<int32_t>(line - (L.side ? d.col0 : d.row0));
  const int32_t* p = L.ptr + (L.side ? d.trp0 : d.arp0);
  base = d.nnz0;
  k0 = p[local];
  k1 = p[local + 1];
  other0 = L.side ? d.row0 : d.col0;
}

// mode 0: largest |a| of each line into stat; 1: Pock-Chambolle sum of |a| in storage order.
__global__ __launch_bounds__(kThreads) void k_bsc_stat(BatchLines L, int mode, double* stat) {
  __shared__ double stage[kThreads * kScU];
  const int lane = threadIdx.x & 63;
  const int64_t line = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (line >= L.lines) return;  // the whole wave leaves together
  int32_t k0, k1, other0;
  int64_t base;
  batch_line(L, line, k0, k1, base, other0);
  const double* v = L.val + base;
  const double r = mode == 0 ? wave_absmax(v, k0, k1, lane)
                             : wave_abssum_ordered(v, k0, k1, lane,
                                                   stage + (threadIdx.x >> 6) * 64 * kScU);
  if (lane == 0) stat[line] = r;
}

// Every entry of the line becomes (a * r_i) * c_j: rows scale by their own factor first,
// columns by the row's (the order of the reference's pass, src/la/scaling.cpp:14-40).
__global__ __launch_bounds__(kThreads) void k_bsc_apply(BatchLines L, const double* line_pass,
                                                        const double* other_pass) {
  const int lane = threadIdx.x & 63;
  const int64_t line = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (line >= L.lines) return;
  int32_t k0, k1, other0;
  int64_t base;
  batch_line(L, line, k0, k1, base, other0);
  wave_apply(L.idx + base, L.val + base, k0, k1, lane, line_pass[line], other_pass + other0,
             L.side == 0 ? 1 : 0);
}

__global__ __launch_bounds__(kThreads) void k_bsc_narrow(const double* v, float* f, int64_t count) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; k < count;
       k += static_cast<int64_t>(gridDim.x) * kThreads)
    f[k] = static_cast<float>(v[k]);
}

// Exact FP32 -> FP64 widening of every LP's iterate pairs (the mixed-precision handoff):
// x = xrest + dx for each half of each pair.
__global__ void k_batch_widen(const float* xs, const double* xr, double* xd, int64_t nx,
                              const float* ys, const double* yr, double* yd, int64_t ny) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nx; k += stride)
    xd[k] = xr[k >> 1] + static_cast<double>(xs[k]);
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ny; k += stride)
    yd[k] = yr[k >> 1] + static_cast<double>(ys[k]);
}

// ---- the cooperative schedule ------------------------------------------------------------
// Every cooperative LP's work items (geometry 0 plans: 256 threads, tiles of 2048) form one
// grid per product, so each attempt of every such LP is three launches shared by all of them:
// the block-diagonal form of the single-LP engine. Per-LP control blocks, per-LP tails.

template <typename V, typename X>
struct CoopArgs {
  Args<V, X> a;
  const int32_t* col_lp;     // LP of each concatenated column
  const int32_t* row_lp;     // LP of each concatenated row
  const int32_t* coop;       // per LP: 1 when it runs in the cooperative schedule
  const CoopItem* act_a;     // the geometry-0 items of Ahat of the LPs still running (grid)
  const CoopItem* act_t;     // ... of Ahat^T
  const int32_t* act_lp;     // the LPs still running (grid of the per-LP kernels)
  double *part_p, *part_d, *part_er, *part_ec, *lr_p;
};

using CG = Geo0;
constexpr int kCT = CG::kThreads;

template <typename V, typename X>
__global__ __launch_bounds__(kCT) void k_coop_primal(CoopArgs<V, X> C, int force) {
  __shared__ double lds[CG::kTile];
  __shared__ double red[4 * 2];
  const CoopItem r = C.act_t[blockIdx.x];
  const int b = r.b, lp = r.lp;
  const Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || (!force && ctl->halt)) return;
  const PlanView<V> T = item_view(C.a.T, C.a.pt, r);
  const Iter<X> it = lp_iter(C.a.it, r.row0, r.col0, r.m, r.n);
  const int4 item = r.item;
  PrimalGather<X> g{it.ybuf, ctl->cur};
  PrimalEpi<X> e = primal_epi(it, ctl);
  Acc<1, 0> acc;
  acc.zero();
  run_item_geo<CG, V, 1>(T, item, g, e, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) C.part_p[b] = acc.s[0];
}

// Long rows of Ahat^T (dense columns) of each cooperative LP, before the dual step reads x'.
template <typename V, typename X>
__global__ __launch_bounds__(kCT) void k_coop_primal_tail(CoopArgs<V, X> C) {
  __shared__ double lds[2 * kChunk];
  __shared__ double red[4];
  const int lp = C.act_lp[blockIdx.x];
  const Ctl* ctl = C.a.ctl + lp;
  const LpDesc& d = C.a.desc[lp];
  if (!C.coop[lp] || ctl->halt || d.t[0].nlr == 0) return;
  const PlanView<V> T = lp_view(C.a.T, C.a.pt, d.t[0], d.nnz0, d.trp0);
  const Iter<X> it = lp_iter(C.a.it, d);
  PrimalEpi<X> e = primal_epi(it, ctl);
  Acc<1, 0> acc;
  acc.zero();
  finish_long_rows_geo<kCT, V, 1>(T, e, acc, lds);
  block_reduce(acc, red);
  if (threadIdx.x == 0) C.lr_p[lp] = acc.s[0];
}

template <typename V, typename X>
__global__ __launch_bounds__(kCT) void k_coop_dual(CoopArgs<V, X> C, int force) {
  __shared__ double lds[CG::kTile];
  __shared__ double red[4 * 2];
  const CoopItem r = C.act_a[blockIdx.x];
  const int b = r.b, lp = r.lp;
  const Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || (!force && ctl->halt)) return;
  const PlanView<V> A = item_view(C.a.A, C.a.pa, r);
  const Iter<X> it = lp_iter(C.a.it, r.row0, r.col0, r.m, r.n);
  const int4 item = r.item;
  DualGather<X> g{it.xbuf, ctl->cur};
  DualEpi<X> e = dual_epi(it, ctl);
  Acc<2, 0> acc;
  acc.zero();
  run_item_geo<CG, V, 2>(A, item, g, e, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    C.part_d[2 * b] = acc.s[0];
    C.part_d[2 * b + 1] = acc.s[1];
  }
}

// Per LP: its long rows of Ahat, its partial sums in index order, and its step decision
// (the single-LP engine's k_dual_tail, one workgroup per LP).
template <typename V, typename X>
__global__ __launch_bounds__(kCT) void k_coop_tail(CoopArgs<V, X> C) {
  __shared__ double lds[2 * kChunk];
  __shared__ double red[4 * 5];
  const int lp = C.act_lp[blockIdx.x];
  Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || ctl->halt) return;
  const LpDesc& d = C.a.desc[lp];
  const PlanView<V> A = lp_view(C.a.A, C.a.pa, d.a[0], d.nnz0, d.arp0);
  const Iter<X> it = lp_iter(C.a.it, d);
  DualEpi<X> epi = dual_epi(it, ctl);
  StepSnap snap{};
  if (threadIdx.x == 0) snap = step_snap(ctl);
  Acc<5, 0> a;
  a.zero();
  const double* pp = C.part_p + d.t[0].item0;
  const double* pd = C.part_d + 2 * static_cast<int64_t>(d.a[0].item0);
  const int np = d.t[0].nitems, nd = d.a[0].nitems;
  const bool has_lr_p = d.t[0].nlr > 0;
  auto partials = [&]() {
    double p[1] = {0.0}, q[2] = {0.0, 0.0};
    lane_sums_1_2(pp, np, pd, nd, p, q);
    if (has_lr_p && threadIdx.x == 0) p[0] += C.lr_p[lp];
    a.s[0] = p[0];
    a.s[1] = q[0];
    a.s[2] = q[1];
    return true;
  };
  Acc<2, 0> lacc;
  lacc.zero();
  finish_long_rows_geo<kCT, V, 2>(A, epi, lacc, lds, partials);
  a.s[3] = lacc.s[0];
  a.s[4] = lacc.s[1];
  block_reduce(a, red);
  if (threadIdx.x == 0) step_decide(ctl, snap, a.s[0], a.s[1] + a.s[3], a.s[2] + a.s[4]);
}

template <typename X>
__global__ __launch_bounds__(kCT) void k_coop_eval_rows(CoopArgs<double, X> C) {
  __shared__ double lds[CG::kTile];
  __shared__ double red[4 * 9];
  const CoopItem r = C.act_a[blockIdx.x];
  const int b = r.b, lp = r.lp;
  const Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || !ctl->tick) return;
  const PlanView<double> A64 = item_view(C.a.A64, C.a.pa, r);
  const Iter<X> it = lp_iter(C.a.it, r.row0, r.col0, r.m, r.n);
  const int4 item = r.item;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickRowsGather<X> g{it, cur, tp, count};
  TickRowsEpi<X> e{it, cur, tp, count};
  Acc<7, 2> acc;
  acc.zero();
  run_item_geo<CG, double, 2>(A64, item, g, e, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) C.part_er[9 * b + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) C.part_er[9 * b + 7 + k] = acc.m[k];
  }
}

template <typename X>
__global__ __launch_bounds__(kCT) void k_coop_eval_cols(CoopArgs<double, X> C) {
  __shared__ double lds[CG::kTile];
  __shared__ double red[4 * 9];
  const CoopItem r = C.act_t[blockIdx.x];
  const int b = r.b, lp = r.lp;
  const Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || !ctl->tick) return;
  const PlanView<double> T64 = item_view(C.a.T64, C.a.pt, r);
  const Iter<X> it = lp_iter(C.a.it, r.row0, r.col0, r.m, r.n);
  const int4 item = r.item;
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickColsGather<X> g{it, cur, tp, count};
  TickColsEpi<X> e{it, cur, tp, count};
  Acc<7, 2> acc;
  acc.zero();
  run_item_geo<CG, double, 2>(T64, item, g, e, acc, lds, red);
  block_reduce(acc, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 7; ++k) C.part_ec[9 * b + k] = acc.s[k];
    for (int k = 0; k < 2; ++k) C.part_ec[9 * b + 7 + k] = acc.m[k];
  }
}

// Per LP: long rows of both products, the tick's totals in index order, the decision.
template <typename X>
__global__ __launch_bounds__(kCT) void k_coop_decide(CoopArgs<double, X> C) {
  __shared__ double lds[2 * kChunk];
  __shared__ double red[4 * 9];
  const int lp = C.act_lp[blockIdx.x];
  Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || !ctl->tick) return;
  const LpDesc& d = C.a.desc[lp];
  const PlanView<double> A64 = lp_view(C.a.A64, C.a.pa, d.a[0], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(C.a.T64, C.a.pt, d.t[0], d.nnz0, d.trp0);
  const Iter<X> it = lp_iter(C.a.it, d);
  const int cur = ctl->cur, tp = ctl->pending;
  const double count = tick_count(ctl);
  TickRowsEpi<X> repi{it, cur, tp, count};
  TickColsEpi<X> cepi{it, cur, tp, count};
  Acc<7, 2> ra, ca;
  ra.zero();
  ca.zero();
  finish_long_rows_geo<kCT, double, 2>(A64, repi, ra, lds);
  finish_long_rows_geo<kCT, double, 2>(T64, cepi, ca, lds);
  block_reduce(ra, red);
  block_reduce(ca, red);
  constexpr bool kMax9[9] = {false, false, false, false, false, false, false, true, true};
  double vr[9], vc[9];
  block_sums<9>(C.part_er + 9 * static_cast<int64_t>(d.a[0].item0), d.a[0].nitems, nullptr, 0, 9,
                kMax9, vr, red);
  block_sums<9>(C.part_ec + 9 * static_cast<int64_t>(d.t[0].item0), d.t[0].nitems, nullptr, 0, 9,
                kMax9, vc, red);
  if (threadIdx.x != 0) return;
  merge9(vr, ra);
  merge9(vc, ca);
  tick_decide(ctl, vr, vc, nullptr);
}

// Apply step over every concatenated coordinate, each under its own LP's decision.
template <typename X>
__global__ __launch_bounds__(kCT) void k_coop_apply(CoopArgs<double, X> C) {
  const Iter<X>& it = C.a.it;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kCT;
  for (int64_t j = blockIdx.x * (int64_t)kCT + threadIdx.x; j < it.n; j += stride) {
    const int lp = C.col_lp[j];
    const Ctl* ctl = C.a.ctl + lp;
    if (!C.coop[lp] || !ctl->apply) continue;
    apply_col(it, apply_args(ctl), j);
  }
  for (int64_t i = blockIdx.x * (int64_t)kCT + threadIdx.x; i < it.m; i += stride) {
    const int lp = C.row_lp[i];
    const Ctl* ctl = C.a.ctl + lp;
    if (!C.coop[lp] || !ctl->apply) continue;
    apply_row(it, apply_args(ctl), i);
  }
}

// Delta-form anchors of each cooperative LP that restarted at this tick (or of every
// cooperative LP when forced, at start), then the end of the tick.
template <typename X>
__global__ __launch_bounds__(kCT) void k_coop_anchor(CoopArgs<double, X> C, int force) {
  __shared__ double lds[lds_doubles<CG>()];
  __shared__ double red[4 * 9];
  const int lp = C.act_lp[blockIdx.x];
  const Ctl* ctl = C.a.ctl + lp;
  if (!C.coop[lp] || (!force && !ctl->restart_now)) return;
  const LpDesc& d = C.a.desc[lp];
  const PlanView<double> A64 = lp_view(C.a.A64, C.a.pa, d.a[0], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(C.a.T64, C.a.pt, d.t[0], d.nnz0, d.trp0);
  anchors<CG>(A64, T64, lp_iter(C.a.it, d), C.a.axr + d.row0, C.a.atyr + d.col0, lds, red);
}

__global__ void k_coop_clear(Ctl* ctl, const int32_t* coop, int nlp) {
  const int lp = blockIdx.x * blockDim.x + threadIdx.x;
  if (lp >= nlp || !coop[lp]) return;
  ctl[lp].apply = 0;
  ctl[lp].copy_best = 0;
  ctl[lp].restart_now = 0;
}

// Start-of-run anchors (delta form) for each LP of the bin.
template <class GEO, int G, typename X>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_anchor(Args<float, X> B) {
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(GEO::kThreads / 64) * 9];
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  const PlanView<double> A64 = lp_view(B.A64, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(B.T64, B.pt, d.t[G], d.nnz0, d.trp0);
  anchors<GEO>(A64, T64, lp_iter(B.it, d), B.axr + d.row0, B.atyr + d.col0, lds, red);
}

// Power iteration of each LP (src/la/scaling.cpp:176-193): v <- Ahat^T Ahat (v / len),
// len = ||v||, norm = ||v||^(1/2); starts from ctl->pw_len, leaves pw_norm / pw_zero.
template <class GEO, int G>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_power(Args<double, double> B, double* v,
                                                              double* u, int iterations) {
  constexpr int NT = GEO::kThreads;
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(NT / 64) * 9];
  __shared__ double len_s, norm_s;
  __shared__ int zero_s;
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  Ctl* ctl = B.ctl + lp;
  const PlanView<double> A64 = lp_view(B.A64, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(B.T64, B.pt, d.t[G], d.nnz0, d.trp0);
  double* vl = v + d.col0;
  double* ul = u + d.row0;
  if (threadIdx.x == 0) {
    len_s = ctl->pw_len;
    norm_s = 0.0;
    zero_s = ctl->pw_zero;
  }
  __syncthreads();
  for (int step = 0; step < iterations && !zero_s; ++step) {
    Acc<1, 0> a1;
    a1.zero();
    PlainGather g1{vl, 0.0, len_s};
    PlainEpi e1{ul};
    walk<GEO, double, 1>(A64, g1, e1, a1, lds, red);
    Acc<1, 0> a2;
    a2.zero();
    PlainGather g2{ul, 0.0, 0.0};
    PlainEpi e2{vl};
    walk<GEO, double, 1>(T64, g2, e2, a2, lds, red);
    block_reduce(a2, red);
    if (threadIdx.x == 0) {
      const double next = a2.s[0];
      norm_s = sqrt(sqrt(next));
      len_s = sqrt(next);
      if (next == 0.0) zero_s = 1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ctl->pw_norm = norm_s;
    ctl->pw_len = len_s;
    ctl->pw_zero = zero_s;
  }
}

// ---- the adjoint ------------------------------------------------------------------------
// Implicit differentiation of the PDHG fixed point. At the solution the projections'
// Jacobians are 0/1 masks: free columns F (strictly inside their bounds) and active rows R
// (equalities, and rows whose multiplier is nonzero, i.e. at a bound). Linearised on that active
// set the fixed point is Ahat_RF xhat_F = bhat_R - Ahat_RN xhat_N and Ahat_RF^T yhat_R = -chat_F,
// so the vector-Jacobian products of the solution map are two least-squares solves with the
// masked matrix M = Ahat_RF, done by CGLS on the same two products PDHG uses, swapped:
//   x side: M^T lambda = Dc gx_F            ->  dL/db_R = Dr lambda
//   y side: M mu = h_R, h = Dr gy          ->  dL/dc_F (minimise space) = -Dc mu
// (gx = dL/dx, gy = dL/dy for the PDHG multiplier y; least squares gives the minimum-norm
// answer where the active set is degenerate). Everything is FP64 and in scaled space.

// Masked product epilogue: out[i] = mask[i] * sum, plus the sum of squares written.
struct MaskEpi {
  double* out;
  const double* mask;
  struct In {
    double k;
  };
  __device__ In load(int i) const { return In{mask[i]}; }
  __device__ void operator()(int i, const RowSum<1>& s, const In& in, Acc<1, 0>& a) const {
    const double v = in.k * s.v[0];
    out[i] = v;
    a.s[0] += v * v;
  }
};

// Reduced cost of the reported point: d_j = cost_j + (A^T y)_j = cost_j + (Ahat^T yhat)_j / Dc_j.
struct RedCostEpi {
  double* out;
  const double *cost, *dc;
  struct In {
    double c, dc;
  };
  __device__ In load(int j) const { return In{cost[j], dc[j]}; }
  __device__ void operator()(int j, const RowSum<1>& s, const In& in, Acc<1, 0>&) const {
    out[j] = in.c + s.v[0] / in.dc;
  }
};

struct AdjWork {
  double *mask_c, *mask_r;      // active set (sum n, sum m)
  const double *gx, *gy;        // upstream gradients (sum n, sum m), unscaled
  double *gb, *gc;              // results (sum m, sum n), unscaled; gc in minimise space
  double *vn0, *vn1, *vn2;      // n-sized CGLS vectors
  double *vm0, *vm1, *vm2;      // m-sized CGLS vectors
  int32_t* iters;               // 2 per LP: CGLS iterations of the x and y solves
  double* resid;                // 2 per LP: final relative normal-equation residuals
};

// One product of the masked matrix: out = mask_out * (P in), returns ||out||^2 (all lanes).
template <class GEO>
__device__ double masked_product(const PlanView<double>& P, const double* in, double* out,
                                 const double* mask_out, double* lds, double* red,
                                 double* bcast) {
  PlainGather g{in, 0.0, 0.0};
  MaskEpi e{out, mask_out};
  Acc<1, 0> a;
  a.zero();
  walk<GEO, double, 1>(P, g, e, a, lds, red);
  block_reduce(a, red);
  if (threadIdx.x == 0) *bcast = a.s[0];
  __syncthreads();
  return *bcast;
}

// CGLS for min || B z - g ||: B = Ahat_RF^T (P1 = Ahat^T plan, P2 = Ahat plan) or B = Ahat_RF
// (the two plans swapped). z and p live on B's columns, r and q on its rows. Returns the
// iterations; *rel receives ||B^T r|| / ||B^T g||.
template <class GEO>
__device__ int cgls(const PlanView<double>& PB, const PlanView<double>& PBt,
                    const double* mask_rows_B, const double* mask_cols_B, int rows_B, int cols_B,
                    double* z, double* r, double* q, double* p, double* sv, double tol,
                    int maxit, double* lds, double* red, double* bcast, double* rel) {
  constexpr int NT = GEO::kThreads;
  // r = g (given), z = 0, s = B^T r, p = s.
  for (int k = threadIdx.x; k < cols_B; k += NT) z[k] = 0.0;
  __syncthreads();
  double gamma = masked_product<GEO>(PBt, r, sv, mask_cols_B, lds, red, bcast);
  const double gamma0 = gamma;
  for (int k = threadIdx.x; k < cols_B; k += NT) p[k] = sv[k];
  __syncthreads();
  int it = 0;
  double g_now = gamma;
  while (it < maxit && gamma0 > 0.0 && g_now > tol * tol * gamma0) {
    const double delta = masked_product<GEO>(PB, p, q, mask_rows_B, lds, red, bcast);
    if (!(delta > 0.0)) break;
    const double alpha = gamma / delta;
    for (int k = threadIdx.x; k < cols_B; k += NT) z[k] += alpha * p[k];
    for (int k = threadIdx.x; k < rows_B; k += NT) r[k] -= alpha * q[k];
    __syncthreads();
    g_now = masked_product<GEO>(PBt, r, sv, mask_cols_B, lds, red, bcast);
    const double beta = g_now / gamma;
    gamma = g_now;
    for (int k = threadIdx.x; k < cols_B; k += NT) p[k] = sv[k] + beta * p[k];
    __syncthreads();
    ++it;
  }
  *rel = gamma0 > 0.0 ? sqrt(g_now / gamma0) : 0.0;
  return it;
}

template <class GEO, int G>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_adjoint(Args<double, double> B,
                                                                AdjWork W, double tol,
                                                                int maxit) {
  constexpr int NT = GEO::kThreads;
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(NT / 64) * 9];
  __shared__ double bcast;
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  const PlanView<double> A64 = lp_view(B.A64, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(B.T64, B.pt, d.t[G], d.nnz0, d.trp0);
  const Iter<double> it = lp_iter(B.it, d);
  const int n = d.n, m = d.m;
  double* mc = W.mask_c + d.col0;
  double* mr = W.mask_r + d.row0;
  // Active set from the reported (unscaled) best point, by complementarity: a column is free
  // when its distance to its nearest bound exceeds its reduced cost |c + A^T y|, and a row is
  // active when it is an equality or its multiplier exceeds its distance to its nearest bound.
  double* m0 = W.vm0 + d.row0;
  double* n0 = W.vn0 + d.col0;
  for (int i = threadIdx.x; i < m; i += NT) m0[i] = it.besty[i] / it.dr[i];
  __syncthreads();
  {
    PlainGather g{m0, 0.0, 0.0};
    RedCostEpi e{n0, it.cost, it.dc};
    Acc<1, 0> a;
    a.zero();
    walk<GEO, double, 1>(T64, g, e, a, lds, red);
  }
  for (int j = threadIdx.x; j < n; j += NT) {
    const double x = it.bestx[j], lo = it.cl[j], hi = it.cu[j];
    double slack = __builtin_huge_val();
    if (finite_bound(lo)) slack = smin(slack, x - lo);
    if (finite_bound(hi)) slack = smin(slack, hi - x);
    mc[j] = slack > fabs(n0[j]) ? 1.0 : 0.0;
  }
  {
    // Row activities for the row slacks: act = A x = Dr^-1 Ahat (x / Dc).
    double* n1 = W.vn1 + d.col0;
    for (int j = threadIdx.x; j < n; j += NT) n1[j] = it.bestx[j] / it.dc[j];
    __syncthreads();
    PlainGather g{n1, 0.0, 0.0};
    PlainEpi e{m0};
    Acc<1, 0> a;
    a.zero();
    walk<GEO, double, 1>(A64, g, e, a, lds, red);
  }
  for (int i = threadIdx.x; i < m; i += NT) {
    const double act = m0[i] / it.dr[i], lo = it.rl[i], hi = it.ru[i];
    double slack = __builtin_huge_val();
    if (finite_bound(lo)) slack = smin(slack, act - lo);
    if (finite_bound(hi)) slack = smin(slack, hi - act);
    mr[i] = (lo == hi || fabs(it.besty[i]) > slack) ? 1.0 : 0.0;
  }
  __syncthreads();
  double* gb = W.gb + d.row0;
  double* gc = W.gc + d.col0;
  for (int i = threadIdx.x; i < m; i += NT) gb[i] = 0.0;
  for (int j = threadIdx.x; j < n; j += NT) gc[j] = 0.0;
  double* n1 = W.vn1 + d.col0;
  double* n2 = W.vn2 + d.col0;
  double* m1 = W.vm1 + d.row0;
  double* m2 = W.vm2 + d.row0;
  // x side: B = M^T (rows: F columns of Ahat, cols: R rows), g = Dc gx on F.
  for (int j = threadIdx.x; j < n; j += NT) n0[j] = mc[j] * it.dc[j] * W.gx[d.col0 + j];
  __syncthreads();
  double rel = 0.0;
  // z = lambda (m0), r = n0, q = n1, p = m1, s = m2; B products: q = T (p), s = A (r).
  int itx = cgls<GEO>(T64, A64, mc, mr, n, m, m0, n0, n1, m1, m2, tol, maxit, lds, red, &bcast,
                      &rel);
  for (int i = threadIdx.x; i < m; i += NT) gb[i] = it.dr[i] * m0[i];
  if (threadIdx.x == 0) {
    W.iters[2 * lp] = itx;
    W.resid[2 * lp] = rel;
  }
  __syncthreads();
  // y side: B = M (rows: R rows, cols: F columns), g = Dr gy on R.
  for (int i = threadIdx.x; i < m; i += NT) m0[i] = mr[i] * it.dr[i] * W.gy[d.row0 + i];
  __syncthreads();
  // z = mu (n0), r = m0, q = m1, p = n1, s = n2; B products: q = A (p), s = T (r).
  const int ity = cgls<GEO>(A64, T64, mr, mc, m, n, n0, m0, m1, n1, n2, tol, maxit, lds, red,
                            &bcast, &rel);
  for (int j = threadIdx.x; j < n; j += NT) gc[j] = -it.dc[j] * n0[j];
  if (threadIdx.x == 0) {
    W.iters[2 * lp + 1] = ity;
    W.resid[2 * lp + 1] = rel;
  }
}

// Report (src/pdhg/pdhg.cpp:486-501) for each LP: evaluate its best point on FP64 values.
template <class GEO, int G, typename X>
__global__ __launch_bounds__(GEO::kThreads) void k_batch_finish(Args<double, X> B,
                                                               double* reduced) {
  constexpr int NT = GEO::kThreads;
  __shared__ double lds[lds_doubles<GEO>()];
  __shared__ double red[(NT / 64) * 9];
  const int lp = B.order[blockIdx.x];
  const LpDesc& d = B.desc[lp];
  Ctl* ctl = B.ctl + lp;
  const PlanView<double> A64 = lp_view(B.A64, B.pa, d.a[G], d.nnz0, d.arp0);
  const PlanView<double> T64 = lp_view(B.T64, B.pt, d.t[G], d.nnz0, d.trp0);
  const Iter<X> it = lp_iter(B.it, d);
  SingleRowsGather<X> gr{it, it.bestx};
  SingleRowsEpi<X> er{it, it.besty, nullptr};
  Acc<3, 1> ra;
  ra.zero();
  walk<GEO, double, 1>(A64, gr, er, ra, lds, red);
  SingleColsGather<X> gc{it, it.besty};
  SingleColsEpi<X> ec{it, it.bestx, reduced + d.col0};
  Acc<3, 1> ca;
  ca.zero();
  walk<GEO, double, 1>(T64, gc, ec, ca, lds, red);
  block_reduce(ra, red);
  block_reduce(ca, red);
  if (threadIdx.x != 0) return;
  // rows: 0 pv 1 sup 2 dv 3 cm ; cols: 0 dv 1 bc 2 po 3 cm
  const Res r = make_res(ra.s[0], ca.s[0] + ra.s[2], smax(ra.m[0], ca.m[0]), ca.s[2],
                         ca.s[1] - ra.s[1], ctl->bound_norm, ctl->cost_norm);
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

}  // namespace batch
}  // namespace sankhya_hip

using namespace sankhya_hip;
using namespace sankhya_hip::batch;

struct pdhg_batch {
  int device = 0;
  int precision = PDHG_PRECISION_FP64;
  int arith = PDHG_PRECISION_FP64;  // the iterates' current precision (mixed: FP32 until promoted)
  hipStream_t stream = nullptr;
  hipStream_t aux[kGeos] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kGeos] = {nullptr, nullptr, nullptr};
  std::string err;
  int32_t nlp = 0;
  int64_t sum_m = 0, sum_n = 0, sum_nnz = 0;
  int64_t padded_total = 0;  // entries of the concatenated (per-LP even-aligned) value arrays
  std::vector<int64_t> lp_nnz;  // nonzeros of each LP
  std::vector<LpDesc> desc_host;
  std::vector<int32_t> geo;                 // per-LP geometry
  std::array<std::vector<int32_t>, kGeos> order;  // per geometry: its LPs, largest first
  LpDesc* desc = nullptr;
  int32_t* order_dev[kGeos] = {nullptr, nullptr, nullptr};
  int32_t* order_all = nullptr;  // every LP, largest first: the one-off per-LP kernels
  int32_t *a_rowptr = nullptr, *a_col = nullptr, *t_rowptr = nullptr, *t_col = nullptr;
  double *a_val64 = nullptr, *t_val64 = nullptr;
  float *a_val32 = nullptr, *t_val32 = nullptr;
  int4 *pa_items[kGeos] = {}, *pa_lr[kGeos] = {}, *pa_grp[kGeos] = {};
  int4 *pt_items[kGeos] = {}, *pt_lr[kGeos] = {}, *pt_grp[kGeos] = {};
  double *pa_seg[kGeos] = {}, *pt_seg[kGeos] = {};
  float *xb32 = nullptr, *yb32 = nullptr;
  double *xb64 = nullptr, *yb64 = nullptr;
  double *xsum = nullptr, *ysum = nullptr, *xrest = nullptr, *yrest = nullptr, *bestx = nullptr,
         *besty = nullptr, *axr = nullptr, *atyr = nullptr, *sx = nullptr, *sy = nullptr;
  double *chat = nullptr, *lhat = nullptr, *uhat = nullptr, *rlhat = nullptr, *ruhat = nullptr;
  double *dc = nullptr, *dr = nullptr, *cost = nullptr, *cl = nullptr, *cu = nullptr, *rl = nullptr,
         *ru = nullptr;
  Ctl* ctl = nullptr;
  Ctl* ctl_host = nullptr;
  long long* clocks = nullptr;
  // Cooperative schedule: LP of every geometry-0 item and coordinate, per-LP mode, partials.
  int32_t *col_lp = nullptr, *row_lp = nullptr;
  int32_t* coop_dev = nullptr;
  std::vector<int32_t> coop;  // per LP: 1 cooperative, 0 persistent (geometry in geo)
  int32_t items_a = 0, items_t = 0;
  // Compacted grids: only the LPs still running (rebuilt as LPs stop), and the host copies of
  // the geometry-0 items the grids' records are made from.
  CoopItem *act_a = nullptr, *act_t = nullptr;
  int32_t* act_lp = nullptr;
  std::vector<int4> items0_a, items0_t;
  int32_t n_act_a = 0, n_act_t = 0, n_act_lp = 0, captured_running = 0;
  bool any_t_long = false;
  double *part_p = nullptr, *part_d = nullptr, *part_er = nullptr, *part_ec = nullptr,
         *lr_p = nullptr;
  // Adjoint work space (allocated on first use).
  double *adj_mc = nullptr, *adj_mr = nullptr, *adj_gx = nullptr, *adj_gy = nullptr,
         *adj_gb = nullptr, *adj_gc = nullptr, *adj_n[3] = {}, *adj_m[3] = {}, *adj_res = nullptr;
  int32_t* adj_it = nullptr;
  hipGraphExec_t exec = nullptr;
  int chunk = kDefaultChunkAttempts;
  bool uploaded = false, started = false;
  int64_t bytes = 0;
  std::vector<void*> allocs;
};

namespace {

int bfail(pdhg_batch* h, int code, const std::string& what) {
  if (h) h->err = what;
  return code;
}

#define BHIP_TRY(h, expr)                                                                  \
  do {                                                                                     \
    const hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                                  \
      return bfail(h, PDHG_E_LAUNCH, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int balloc(pdhg_batch* h, T** p, size_t count) {
  *p = nullptr;
  // kSpare zeroed elements past the end: a lane outside its item gathers index 0 of its LP, which
  // for an LP with no rows (or no columns) is the next LP's first entry, or, for the last LP of
  // the batch, one past the array.
  constexpr size_t kSpare = 2;
  const size_t bytes = (count + kSpare) * sizeof(T);
  if (hipMalloc(reinterpret_cast<void**>(p), bytes) != hipSuccess) {
    *p = nullptr;
    return bfail(h, PDHG_E_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  h->allocs.push_back(*p);
  h->bytes += static_cast<int64_t>(bytes);
  BHIP_TRY(h, hipMemsetAsync(*p + count, 0, kSpare * sizeof(T), h->stream));
  return PDHG_OK;
}

template <typename T>
int bupload(pdhg_batch* h, T** dst, const T* src, size_t count) {
  int rc = balloc(h, dst, count);
  if (rc) return rc;
  if (count && src)
    BHIP_TRY(h, hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

// True while the iterates are FP32 (a mixed batch before pdhg_batch_promote).
bool mixed(const pdhg_batch* h) { return h->arith == PDHG_PRECISION_MIXED; }

template <typename V>
Mat<V> mat(const int32_t* rp, const int32_t* col, const V* val) {
  return Mat<V>{rp, col, val};
}

template <typename V, typename X>
Args<V, X> args_of(const pdhg_batch* h, int g) {
  Args<V, X> a;
  a.desc = h->desc;
  a.order = h->order_dev[g];
  a.ctl = h->ctl;
  if constexpr (sizeof(V) == 8) {
    a.A = mat<V>(h->a_rowptr, h->a_col, h->a_val64);
    a.T = mat<V>(h->t_rowptr, h->t_col, h->t_val64);
  } else {
    a.A = mat<V>(h->a_rowptr, h->a_col, h->a_val32);
    a.T = mat<V>(h->t_rowptr, h->t_col, h->t_val32);
  }
  a.A64 = mat<double>(h->a_rowptr, h->a_col, h->a_val64);
  a.T64 = mat<double>(h->t_rowptr, h->t_col, h->t_val64);
  a.pa = PlanArrays{h->pa_items[g], h->pa_lr[g], h->pa_grp[g], h->pa_seg[g]};
  a.pt = PlanArrays{h->pt_items[g], h->pt_lr[g], h->pt_grp[g], h->pt_seg[g]};
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
  it.n = static_cast<int32_t>(h->sum_n);
  it.m = static_cast<int32_t>(h->sum_m);
  a.it = it;
  a.axr = h->axr;
  a.atyr = h->atyr;
  a.clocks = h->clocks;
  return a;
}

// Per-geometry launches.
template <typename V, typename X>
void launch_run(pdhg_batch* h, int g, hipStream_t s, int attempts) {
  const int nb = static_cast<int>(h->order[g].size());
  if (nb == 0) return;
  const Args<V, X> a = args_of<V, X>(h, g);
  if (g == 0) k_batch_run<Geo0, 0, V, X><<<nb, Geo0::kThreads, 0, s>>>(a, attempts);
  else if (g == 1) k_batch_run<Geo1, 1, V, X><<<nb, Geo1::kThreads, 0, s>>>(a, attempts);
  else k_batch_run<Geo2, 2, V, X><<<nb, Geo2::kThreads, 0, s>>>(a, attempts);
}

template <typename X>
void launch_tick(pdhg_batch* h, int g, hipStream_t s) {
  const int nb = static_cast<int>(h->order[g].size());
  if (nb == 0) return;
  const Args<double, X> a = args_of<double, X>(h, g);
  if (g == 0) k_batch_tick<Geo0, 0, X><<<nb, Geo0::kThreads, 0, s>>>(a);
  else if (g == 1) k_batch_tick<Geo1, 1, X><<<nb, Geo1::kThreads, 0, s>>>(a);
  else k_batch_tick<Geo2, 2, X><<<nb, Geo2::kThreads, 0, s>>>(a);
}

// `rounds` x (up to kSlotAttempts attempts, then the tick) for one bin.
void launch_rounds(pdhg_batch* h, int g, hipStream_t s, int rounds) {
  for (int r = 0; r < rounds; ++r) {
    if (mixed(h)) {
      launch_run<float, float>(h, g, s, kSlotAttempts);
      launch_tick<float>(h, g, s);
    } else {
      launch_run<double, double>(h, g, s, kSlotAttempts);
      launch_tick<double>(h, g, s);
    }
  }
}

template <typename V, typename X>
CoopArgs<V, X> coop_args(const pdhg_batch* h) {
  CoopArgs<V, X> c;
  c.a = args_of<V, X>(h, 0);
  c.col_lp = h->col_lp;
  c.row_lp = h->row_lp;
  c.coop = h->coop_dev;
  c.act_a = h->act_a;
  c.act_t = h->act_t;
  c.act_lp = h->act_lp;
  c.part_p = h->part_p;
  c.part_d = h->part_d;
  c.part_er = h->part_er;
  c.part_ec = h->part_ec;
  c.lr_p = h->lr_p;
  return c;
}

bool any_coop(const pdhg_batch* h) {
  return std::any_of(h->coop.begin(), h->coop.end(), [](int32_t c) { return c != 0; });
}

// One cooperative attempt for every cooperative LP (three or four launches).
template <typename V, typename X>
void launch_coop_attempt(pdhg_batch* h, hipStream_t s, int force) {
  const CoopArgs<V, X> c = coop_args<V, X>(h);
  if (h->n_act_lp == 0) return;
  if (h->n_act_t > 0) k_coop_primal<V, X><<<h->n_act_t, kCT, 0, s>>>(c, force);
  if (h->any_t_long) k_coop_primal_tail<V, X><<<h->n_act_lp, kCT, 0, s>>>(c);
  if (h->n_act_a > 0) k_coop_dual<V, X><<<h->n_act_a, kCT, 0, s>>>(c, force);
  k_coop_tail<V, X><<<h->n_act_lp, kCT, 0, s>>>(c);
}

// The cooperative evaluation tick.
template <typename X>
void launch_coop_tick(pdhg_batch* h, hipStream_t s) {
  const CoopArgs<double, X> c = coop_args<double, X>(h);
  if (h->n_act_lp == 0) return;
  if (h->n_act_a > 0) k_coop_eval_rows<X><<<h->n_act_a, kCT, 0, s>>>(c);
  if (h->n_act_t > 0) k_coop_eval_cols<X><<<h->n_act_t, kCT, 0, s>>>(c);
  k_coop_decide<X><<<h->n_act_lp, kCT, 0, s>>>(c);
  const int64_t w = std::max<int64_t>(std::max(h->sum_n, h->sum_m), 1);
  const int grid = static_cast<int>(std::min<int64_t>((w + kCT - 1) / kCT, 8192));
  k_coop_apply<X><<<grid, kCT, 0, s>>>(c);
  if constexpr (kDelta<X>) k_coop_anchor<X><<<h->n_act_lp, kCT, 0, s>>>(c, 0);
  k_coop_clear<<<(h->nlp + 255) / 256, 256, 0, s>>>(h->ctl, h->coop_dev, h->nlp);
}

// `rounds` x (kSlotAttempts cooperative attempts, then the tick).
void launch_coop_rounds(pdhg_batch* h, hipStream_t s, int rounds) {
  for (int r = 0; r < rounds; ++r) {
    for (int a = 0; a < kSlotAttempts; ++a) {
      if (mixed(h)) launch_coop_attempt<float, float>(h, s, 0);
      else launch_coop_attempt<double, double>(h, s, 0);
    }
    if (mixed(h)) launch_coop_tick<float>(h, s);
    else launch_coop_tick<double>(h, s);
  }
}

void launch_anchor(pdhg_batch* h, int g) {
  const int nb = static_cast<int>(h->order[g].size());
  if (nb == 0 || !mixed(h)) return;
  const Args<float, float> a = args_of<float, float>(h, g);
  if (g == 0) k_batch_anchor<Geo0, 0, float><<<nb, Geo0::kThreads, 0, h->stream>>>(a);
  else if (g == 1) k_batch_anchor<Geo1, 1, float><<<nb, Geo1::kThreads, 0, h->stream>>>(a);
  else k_batch_anchor<Geo2, 2, float><<<nb, Geo2::kThreads, 0, h->stream>>>(a);
}

// The one-off per-LP kernels (power iteration, report) run every LP, whatever its schedule,
// one 512-thread workgroup each.
void launch_power(pdhg_batch* h, int iterations) {
  Args<double, double> a = args_of<double, double>(h, 1);
  a.order = h->order_all;
  k_batch_power<Geo1, 1><<<h->nlp, Geo1::kThreads, 0, h->stream>>>(a, h->sx, h->sy, iterations);
}

template <typename X>
void launch_finish(pdhg_batch* h, double* reduced) {
  Args<double, X> a = args_of<double, X>(h, 1);
  a.order = h->order_all;
  k_batch_finish<Geo1, 1, X><<<h->nlp, Geo1::kThreads, 0, h->stream>>>(a, reduced);
}

int compact(pdhg_batch* h, bool all);

// Per-geometry LP lists, largest LP first (it has the longest run, so it starts first).
int set_bins(pdhg_batch* h) {
  for (int g = 0; g < kGeos; ++g) h->order[g].clear();
  std::vector<int32_t> idx(h->nlp);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
    const LpDesc &da = h->desc_host[a], &db = h->desc_host[b];
    const int64_t wa = int64_t{da.a[0].nitems} + da.t[0].nitems + da.m + da.n;
    const int64_t wb = int64_t{db.a[0].nitems} + db.t[0].nitems + db.m + db.n;
    return wa > wb;
  });
  for (const int32_t k : idx)
    if (!h->coop[k]) h->order[h->geo[k]].push_back(k);
  if (!h->order_all) {
    int rc = balloc(h, &h->order_all, static_cast<size_t>(h->nlp));
    if (rc) return rc;
  }
  BHIP_TRY(h, hipMemcpyAsync(h->order_all, idx.data(), sizeof(int32_t) * h->nlp,
                             hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  {
    const int rc = compact(h, true);
    if (rc) return rc;
  }
  if (!h->coop_dev) {
    int rc = balloc(h, &h->coop_dev, static_cast<size_t>(h->nlp));
    if (rc) return rc;
  }
  BHIP_TRY(h, hipMemcpyAsync(h->coop_dev, h->coop.data(), sizeof(int32_t) * h->nlp,
                             hipMemcpyHostToDevice, h->stream));
  for (int g = 0; g < kGeos; ++g) {
    if (!h->order_dev[g]) {
      int rc = balloc(h, &h->order_dev[g], static_cast<size_t>(h->nlp));
      if (rc) return rc;
    }
    if (!h->order[g].empty())
      BHIP_TRY(h, hipMemcpyAsync(h->order_dev[g], h->order[g].data(),
                                 sizeof(int32_t) * h->order[g].size(), hipMemcpyHostToDevice,
                                 h->stream));
  }
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

// The cooperative grids over the cooperative LPs that are still running (`running` from the
// host copy of the control blocks; all of them when `all` is set).
int compact(pdhg_batch* h, bool all) {
  // XCD placement: grid position p runs on XCD p % kXcds, and each XCD has its own L2. Every
  // running LP is given one XCD (largest LP first, to the least loaded), and its items take
  // that XCD's positions, so an LP's iterate vectors are fetched into one L2 instead of all
  // eight. The order of the items does not change any result: partial sums are indexed by
  // item and added per LP in item order.
  constexpr int kXcds = 8;
  std::vector<int32_t> lps;
  for (int32_t k = 0; k < h->nlp; ++k)
    if (h->coop[k] && (all || !h->ctl_host[k].halt)) lps.push_back(k);
  auto work = [&](int32_t k) {
    const LpDesc& d = h->desc_host[k];
    return int64_t{d.a[0].nitems} + d.t[0].nitems;
  };
  std::vector<int32_t> by_size(lps);
  std::stable_sort(by_size.begin(), by_size.end(),
                   [&](int32_t a, int32_t b) { return work(a) > work(b); });
  std::vector<std::vector<int32_t>> lps_of(kXcds);
  std::array<int64_t, kXcds> load{};
  for (const int32_t k : by_size) {
    int x = 0;
    for (int c = 1; c < kXcds; ++c)
      if (load[c] < load[x]) x = c;
    lps_of[x].push_back(k);
    load[x] += work(k);
  }
  // Items of one product: each XCD's queue holds its LPs' items LP after LP; position p takes
  // the next item of queue p % kXcds, or of the fullest queue once that one is empty.
  auto place = [&](bool rows_of_a) {
    std::vector<std::vector<CoopItem>> q(kXcds);
    const std::vector<int4>& items0 = rows_of_a ? h->items0_a : h->items0_t;
    size_t total = 0;
    for (int x = 0; x < kXcds; ++x)
      for (const int32_t k : lps_of[x]) {
        const LpDesc& d = h->desc_host[k];
        const LpPlan& pd = rows_of_a ? d.a[0] : d.t[0];
        for (int32_t i = 0; i < pd.nitems; ++i) {
          CoopItem r{};
          r.item = items0[static_cast<size_t>(pd.item0 + i)];
          r.b = pd.item0 + i;
          r.lp = k;
          r.rp0 = rows_of_a ? d.arp0 : d.trp0;
          r.lr0 = pd.lr0;
          r.nnz0 = d.nnz0;
          r.seg0 = pd.seg0;
          r.row0 = d.row0;
          r.col0 = d.col0;
          r.m = d.m;
          r.n = d.n;
          q[x].push_back(r);
        }
        total += static_cast<size_t>(pd.nitems);
      }
    std::vector<CoopItem> out;
    out.reserve(total);
    std::array<size_t, kXcds> next{};
    for (size_t p = 0; p < total; ++p) {
      int x = static_cast<int>(p % kXcds);
      if (next[x] >= q[x].size()) {
        for (int c = 0; c < kXcds; ++c)
          if (q[c].size() - next[c] > q[x].size() - next[x]) x = c;
      }
      out.push_back(q[x][next[x]++]);
    }
    return out;
  };
  const std::vector<CoopItem> ia = place(true), it = place(false);
  h->n_act_lp = static_cast<int32_t>(lps.size());
  h->n_act_a = static_cast<int32_t>(ia.size());
  h->n_act_t = static_cast<int32_t>(it.size());
  if (!lps.empty())
    BHIP_TRY(h, hipMemcpyAsync(h->act_lp, lps.data(), sizeof(int32_t) * lps.size(), hipMemcpyHostToDevice, h->stream));
  if (!ia.empty())
    BHIP_TRY(h, hipMemcpyAsync(h->act_a, ia.data(), sizeof(CoopItem) * ia.size(), hipMemcpyHostToDevice, h->stream));
  if (!it.empty())
    BHIP_TRY(h, hipMemcpyAsync(h->act_t, it.data(), sizeof(CoopItem) * it.size(), hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

// One graph chunk: each non-empty bin's launch on its own branch, joined at the end.
int capture(pdhg_batch* h) {
  if (h->exec) {
    BHIP_TRY(h, hipGraphExecDestroy(h->exec));
    h->exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  BHIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  BHIP_TRY(h, hipEventRecord(h->fork, h->stream));
  const int coop_rounds = std::max(1, (h->chunk + kSlotAttempts - 1) / kSlotAttempts);
  bool first = true;
  if (any_coop(h)) {
    launch_coop_rounds(h, h->stream, coop_rounds);
    first = false;
  }
  for (int g = kGeos - 1; g >= 0; --g) {
    if (h->order[g].empty()) continue;
    const int rounds = std::max(1, (h->chunk + kSlotAttempts - 1) / kSlotAttempts);
    if (first) {
      launch_rounds(h, g, h->stream, rounds);
      first = false;
      continue;
    }
    BHIP_TRY(h, hipStreamWaitEvent(h->aux[g], h->fork, 0));
    launch_rounds(h, g, h->aux[g], rounds);
    BHIP_TRY(h, hipEventRecord(h->join[g], h->aux[g]));
    BHIP_TRY(h, hipStreamWaitEvent(h->stream, h->join[g], 0));
  }
  BHIP_TRY(h, hipStreamEndCapture(h->stream, &graph));
  const hipError_t e = hipGraphInstantiate(&h->exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess)
    return bfail(h, PDHG_E_LAUNCH, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return PDHG_OK;
}

int check(pdhg_batch* h, bool need_start) {
  if (!h) return PDHG_E_ARGUMENT;
  if (!h->uploaded) return bfail(h, PDHG_E_STATE, "no batch uploaded");
  if (need_start && !h->started) return bfail(h, PDHG_E_STATE, "batch not started");
  return PDHG_OK;
}

// Every LP's preconditioner on the device: `ruiz` Ruiz rounds and one Pock-Chambolle round in
// place in both copies of the values, Dr and Dc, then the scaled cost and bounds, and the FP32
// copies of the values in mixed precision. Scratch: sy and yrest hold the row statistics and
// factors, sx and xrest the column ones (start_state sets every iterate array afterwards).
int batch_device_scale(pdhg_batch* h, int ruiz) {
  const int64_t sm = h->sum_m, sn = h->sum_n;
  const BatchLines rows{h->a_rowptr, h->a_col, h->a_val64, h->desc, h->row_lp, sm, 0};
  const BatchLines cols{h->t_rowptr, h->t_col, h->t_val64, h->desc, h->col_lp, sn, 1};
  double *rstat = h->sy, *rpass = h->yrest, *cstat = h->sx, *cpass = h->xrest;
  auto waves = [](int64_t lines) { return static_cast<unsigned>((lines + kThreads / 64 - 1) / (kThreads / 64)); };
  auto lanes = [](int64_t count) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((count + kThreads - 1) / kThreads, 8192)));
  };
  if (sm > 0) k_sc_fill<<<lanes(sm), kThreads, 0, h->stream>>>(h->dr, static_cast<int32_t>(sm), 1.0);
  if (sn > 0) k_sc_fill<<<lanes(sn), kThreads, 0, h->stream>>>(h->dc, static_cast<int32_t>(sn), 1.0);
  auto round = [&](int mode) {
    if (sm > 0) k_bsc_stat<<<waves(sm), kThreads, 0, h->stream>>>(rows, mode, rstat);
    if (sn > 0) k_bsc_stat<<<waves(sn), kThreads, 0, h->stream>>>(cols, mode, cstat);
    if (sm > 0) k_sc_factor<<<lanes(sm), kThreads, 0, h->stream>>>(rstat, rpass, h->dr, static_cast<int32_t>(sm));
    if (sn > 0) k_sc_factor<<<lanes(sn), kThreads, 0, h->stream>>>(cstat, cpass, h->dc, static_cast<int32_t>(sn));
    if (sm > 0) k_bsc_apply<<<waves(sm), kThreads, 0, h->stream>>>(rows, rpass, cpass);
    if (sn > 0) k_bsc_apply<<<waves(sn), kThreads, 0, h->stream>>>(cols, cpass, rpass);
  };
  for (int pass = 0; pass < ruiz; ++pass) round(0);
  round(1);
  const int64_t most = std::max<int64_t>(std::max(sm, sn), 1);
  k_sc_vectors<<<lanes(most), kThreads, 0, h->stream>>>(
      static_cast<int32_t>(sn), static_cast<int32_t>(sm), h->dc, h->dr, h->cost, h->cl, h->cu, h->rl,
      h->ru, h->chat, h->lhat, h->uhat, h->rlhat, h->ruhat);
  if (h->precision == PDHG_PRECISION_MIXED && h->padded_total > 0) {
    k_bsc_narrow<<<lanes(h->padded_total), kThreads, 0, h->stream>>>(h->a_val64, h->a_val32, h->padded_total);
    k_bsc_narrow<<<lanes(h->padded_total), kThreads, 0, h->stream>>>(h->t_val64, h->t_val32, h->padded_total);
  }
  BHIP_TRY(h, hipGetLastError());
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

// Control blocks at the start point (src/pdhg/pdhg.cpp:137-192), iterates initialised, and the
// delta form's anchors.
int start_state(pdhg_batch* h, const pdhg_params* p, const double* norms) {
  h->arith = h->precision;  // a mixed batch starts (again) in its FP32 phase
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost,
                             h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int32_t k = 0; k < h->nlp; ++k) {
    Ctl& c = h->ctl_host[k];
    const double bn = c.bound_norm, cn = c.cost_norm;
    std::memset(&c, 0, sizeof(c));
    c.bound_norm = bn;
    c.cost_norm = cn;
    const double sn = norms ? norms[k] : 1.0;
    c.eta = sn > 0.0 ? 1.0 / sn : 1.0;
    c.omega = 1.0;
    c.eta_ceiling = 1.0e3 / std::max(sn, 1e-12);
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
  }
  BHIP_TRY(h, hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl) * h->nlp, hipMemcpyHostToDevice,
                             h->stream));
  Iter<double> itd = args_of<double, double>(h, 0).it;
  Iter<float> itf = args_of<float, float>(h, 0).it;
  const int64_t w = std::max<int64_t>(std::max(h->sum_n, h->sum_m), 1);
  const int grid = static_cast<int>(std::min<int64_t>((w + 255) / 256, 8192));
  if (mixed(h)) k_batch_init<float><<<grid, 256, 0, h->stream>>>(itf);
  else k_batch_init<double><<<grid, 256, 0, h->stream>>>(itd);
  BHIP_TRY(h, hipGetLastError());
  for (int g = 0; g < kGeos; ++g) launch_anchor(h, g);
  if (mixed(h) && any_coop(h) && h->n_act_lp > 0)
    k_coop_anchor<float><<<h->n_act_lp, kCT, 0, h->stream>>>(coop_args<double, float>(h), 1);
  BHIP_TRY(h, hipGetLastError());
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

}  // namespace

extern "C" {

int pdhg_batch_create(int device, int precision, pdhg_batch** out) {
  if (!out) return PDHG_E_ARGUMENT;
  *out = nullptr;
  if (precision != PDHG_PRECISION_FP64 && precision != PDHG_PRECISION_MIXED) return PDHG_E_ARGUMENT;
  if (pdhg_dev_probe(device, nullptr, 0) != PDHG_OK) return PDHG_E_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PDHG_E_NO_DEVICE;
  pdhg_batch* h = new pdhg_batch();
  h->device = device;
  h->precision = precision;
  h->arith = precision;
  bool ok = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) == hipSuccess;
  for (int g = 0; ok && g < kGeos; ++g)
    ok = hipStreamCreateWithFlags(&h->aux[g], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&h->join[g], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    pdhg_batch_destroy(h);
    return PDHG_E_LAUNCH;
  }
  *out = h;
  return PDHG_OK;
}

void pdhg_batch_destroy(pdhg_batch* h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->exec) hipGraphExecDestroy(h->exec);
  for (void* p : h->allocs) hipFree(p);
  if (h->ctl_host) hipHostFree(h->ctl_host);
  for (int g = 0; g < kGeos; ++g) {
    if (h->aux[g]) hipStreamDestroy(h->aux[g]);
    if (h->join[g]) hipEventDestroy(h->join[g]);
  }
  if (h->fork) hipEventDestroy(h->fork);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

const char* pdhg_batch_last_error(const pdhg_batch* h) { return h ? h->err.c_str() : "null handle"; }

int64_t pdhg_batch_bytes(const pdhg_batch* h) { return h ? h->bytes : 0; }

int pdhg_batch_upload(pdhg_batch* h, int32_t nlp, const int32_t* m, const int32_t* n,
                      const int64_t* nnz, const int32_t* a_rowptr, const int32_t* a_col,
                      const double* a_val, const int32_t* at_rowptr, const int32_t* at_col,
                      const double* at_val, const double* cost_hat, const double* lower_hat,
                      const double* upper_hat, const double* row_lower_hat,
                      const double* row_upper_hat, const double* col_scale,
                      const double* row_scale, const double* min_cost, const double* col_lower,
                      const double* col_upper, const double* row_lower, const double* row_upper,
                      const double* bound_norm, const double* cost_norm) {
  if (!h) return PDHG_E_ARGUMENT;
  if (h->uploaded) return bfail(h, PDHG_E_STATE, "batch already uploaded");
  if (!bound_norm || !cost_norm) return bfail(h, PDHG_E_ARGUMENT, "null norms");
  BatchLayout L;
  const std::string what = build_batch_layout(nlp, m, n, nnz, a_rowptr, a_col, a_val, at_rowptr,
                                              at_col, at_val, &L);
  if (!what.empty()) return bfail(h, PDHG_E_ARGUMENT, what);
  BHIP_TRY(h, hipSetDevice(h->device));
  h->nlp = nlp;
  h->desc_host = L.desc;
  h->lp_nnz.assign(nnz, nnz + nlp);
  h->geo.assign(static_cast<size_t>(nlp), kDefaultGeo);
  h->sum_m = L.sum_m;
  h->sum_n = L.sum_n;
  h->sum_nnz = L.sum_nnz;
  const int64_t sm = L.sum_m, sn = L.sum_n;
  {
    const size_t total = L.a_col.size();
    h->padded_total = static_cast<int64_t>(total);
    int64_t arp = 0, trp = 0;
    for (int32_t k = 0; k < nlp; ++k) {
      arp += m[k] + 1;
      trp += n[k] + 1;
    }
    int rc;
    if ((rc = bupload(h, &h->a_col, L.a_col.data(), total)) || (rc = bupload(h, &h->t_col, L.t_col.data(), total)) ||
        (rc = bupload(h, &h->a_val64, L.a_val.data(), total)) || (rc = bupload(h, &h->t_val64, L.t_val.data(), total)) ||
        (rc = bupload(h, &h->a_rowptr, a_rowptr, static_cast<size_t>(arp))) ||
        (rc = bupload(h, &h->t_rowptr, at_rowptr, static_cast<size_t>(trp))))
      return rc;
    if (h->precision == PDHG_PRECISION_MIXED) {
      std::vector<float> af(total), tf(total);
      for (size_t q = 0; q < total; ++q) {
        af[q] = static_cast<float>(L.a_val[q]);
        tf[q] = static_cast<float>(L.t_val[q]);
      }
      if ((rc = bupload(h, &h->a_val32, af.data(), total)) || (rc = bupload(h, &h->t_val32, tf.data(), total)))
        return rc;
    }
    for (int g = 0; g < kGeos; ++g) {
      for (int side = 0; side < 2; ++side) {
        const PlanSet& ps = L.plans[side][g];
        int4** it = side == 0 ? &h->pa_items[g] : &h->pt_items[g];
        int4** lp = side == 0 ? &h->pa_lr[g] : &h->pt_lr[g];
        int4** gp = side == 0 ? &h->pa_grp[g] : &h->pt_grp[g];
        double** sp = side == 0 ? &h->pa_seg[g] : &h->pt_seg[g];
        if ((rc = bupload(h, it, ps.items.data(), ps.items.size())) ||
            (rc = bupload(h, lp, ps.lr_info.data(), ps.lr_info.size())) ||
            (rc = bupload(h, gp, ps.groups.data(), ps.groups.size())) ||
            (rc = balloc(h, sp, static_cast<size_t>(2 * ps.segs))))
          return rc;
      }
    }
  }
  int rc;
  if ((rc = bupload(h, &h->desc, h->desc_host.data(), h->desc_host.size()))) return rc;
  const size_t un = static_cast<size_t>(sn), um = static_cast<size_t>(sm);
  if (h->precision == PDHG_PRECISION_MIXED) {
    if ((rc = balloc(h, &h->xb32, 2 * un)) || (rc = balloc(h, &h->yb32, 2 * um))) return rc;
  } else if ((rc = balloc(h, &h->xb64, 2 * un)) || (rc = balloc(h, &h->yb64, 2 * um))) {
    return rc;
  }
  double** nv[] = {&h->xsum, &h->xrest, &h->bestx, &h->atyr, &h->sx};
  for (double** v : nv)
    if ((rc = balloc(h, v, un))) return rc;
  double** mv[] = {&h->ysum, &h->yrest, &h->besty, &h->axr, &h->sy};
  for (double** v : mv)
    if ((rc = balloc(h, v, um))) return rc;
  if ((rc = bupload(h, &h->chat, cost_hat, un)) || (rc = bupload(h, &h->lhat, lower_hat, un)) ||
      (rc = bupload(h, &h->uhat, upper_hat, un)) || (rc = bupload(h, &h->rlhat, row_lower_hat, um)) ||
      (rc = bupload(h, &h->ruhat, row_upper_hat, um)) || (rc = bupload(h, &h->dc, col_scale, un)) ||
      (rc = bupload(h, &h->dr, row_scale, um)) || (rc = bupload(h, &h->cost, min_cost, un)) ||
      (rc = bupload(h, &h->cl, col_lower, un)) || (rc = bupload(h, &h->cu, col_upper, un)) ||
      (rc = bupload(h, &h->rl, row_lower, um)) || (rc = bupload(h, &h->ru, row_upper, um)) ||
      (rc = balloc(h, &h->ctl, static_cast<size_t>(nlp))) ||
      (rc = balloc(h, &h->clocks, static_cast<size_t>(nlp))))
    return rc;
  {
    // The geometry-0 items (the cooperative grids' records are made from them) and the LP of
    // each coordinate.
    const PlanSet& pa0 = L.plans[0][0];
    const PlanSet& pt0 = L.plans[1][0];
    h->items0_a = pa0.items;
    h->items0_t = pt0.items;
    const size_t ia = pa0.items.size(), it = pt0.items.size();
    std::vector<int32_t> cl_lp(un), rl_lp(um);
    for (int32_t k = 0; k < nlp; ++k) {
      const LpDesc& d = L.desc[k];
      for (int32_t j = 0; j < d.n; ++j) cl_lp[d.col0 + j] = k;
      for (int32_t i = 0; i < d.m; ++i) rl_lp[d.row0 + i] = k;
      if (d.t[0].nlr > 0) h->any_t_long = true;
    }
    h->items_a = static_cast<int32_t>(ia);
    h->items_t = static_cast<int32_t>(it);
    if ((rc = bupload(h, &h->col_lp, cl_lp.data(), cl_lp.size())) ||
        (rc = bupload(h, &h->row_lp, rl_lp.data(), rl_lp.size())) ||
        (rc = balloc(h, &h->part_p, it)) || (rc = balloc(h, &h->part_d, 2 * ia)) ||
        (rc = balloc(h, &h->part_er, 9 * ia)) || (rc = balloc(h, &h->part_ec, 9 * it)) ||
        (rc = balloc(h, &h->lr_p, static_cast<size_t>(nlp))) ||
        (rc = balloc(h, &h->act_a, ia)) || (rc = balloc(h, &h->act_t, it)) ||
        (rc = balloc(h, &h->act_lp, static_cast<size_t>(nlp))))
      return rc;
    h->coop.assign(static_cast<size_t>(nlp), 1);
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&h->ctl_host), sizeof(Ctl) * nlp, hipHostMallocDefault) != hipSuccess)
    return bfail(h, PDHG_E_ALLOC, "pinned control blocks");
  std::memset(h->ctl_host, 0, sizeof(Ctl) * nlp);
  for (int32_t k = 0; k < nlp; ++k) {
    h->ctl_host[k].bound_norm = bound_norm[k];
    h->ctl_host[k].cost_norm = cost_norm[k];
    h->ctl_host[k].halt = 1;
  }
  BHIP_TRY(h, hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl) * nlp, hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  if ((rc = set_bins(h))) return rc;
  h->uploaded = true;
  return PDHG_OK;
}

int pdhg_batch_upload_unscaled(pdhg_batch* h, int32_t nlp, const int32_t* m, const int32_t* n,
                               const int64_t* nnz, const int32_t* a_rowptr, const int32_t* a_col,
                               const double* a_val, const int32_t* at_rowptr, const int32_t* at_col,
                               const double* at_val, const double* min_cost,
                               const double* col_lower, const double* col_upper,
                               const double* row_lower, const double* row_upper,
                               const double* bound_norm, const double* cost_norm,
                               int ruiz_iterations) {
  if (!h) return PDHG_E_ARGUMENT;
  if (ruiz_iterations < 0) return bfail(h, PDHG_E_ARGUMENT, "negative Ruiz iteration count");
  // The unscaled vectors stand in for the scaled ones until the device passes overwrite them.
  int rc = pdhg_batch_upload(h, nlp, m, n, nnz, a_rowptr, a_col, a_val, at_rowptr, at_col, at_val,
                             min_cost, col_lower, col_upper, row_lower, row_upper, nullptr,
                             nullptr, min_cost, col_lower, col_upper, row_lower, row_upper,
                             bound_norm, cost_norm);
  if (rc) return rc;
  return batch_device_scale(h, ruiz_iterations);
}

int pdhg_batch_scaling(pdhg_batch* h, double* row, double* column, double* csc_values,
                       double* cost_hat, double* col_lower_hat, double* col_upper_hat,
                       double* row_lower_hat, double* row_upper_hat) {
  int rc = check(h, false);
  if (rc) return rc;
  BHIP_TRY(h, hipSetDevice(h->device));
  const size_t un = static_cast<size_t>(h->sum_n), um = static_cast<size_t>(h->sum_m);
  struct Item {
    double* dst;
    const double* src;
    size_t count;
  } items[] = {{row, h->dr, um},          {column, h->dc, un},          {cost_hat, h->chat, un},
               {col_lower_hat, h->lhat, un}, {col_upper_hat, h->uhat, un}, {row_lower_hat, h->rlhat, um},
               {row_upper_hat, h->ruhat, um}};
  for (const Item& it : items)
    if (it.dst && it.count)
      BHIP_TRY(h, hipMemcpyAsync(it.dst, it.src, sizeof(double) * it.count, hipMemcpyDeviceToHost, h->stream));
  if (csc_values && h->sum_nnz > 0) {
    std::vector<double> all(static_cast<size_t>(h->padded_total));
    BHIP_TRY(h, hipMemcpyAsync(all.data(), h->t_val64, sizeof(double) * all.size(), hipMemcpyDeviceToHost, h->stream));
    BHIP_TRY(h, hipStreamSynchronize(h->stream));
    // Each LP's entries open its (even-aligned) block of the concatenated array.
    int64_t out = 0;
    for (int32_t k = 0; k < h->nlp; ++k) {
      const int64_t count = h->lp_nnz[static_cast<size_t>(k)];
      std::memcpy(csc_values + out, all.data() + h->desc_host[k].nnz0,
                  sizeof(double) * static_cast<size_t>(count));
      out += count;
    }
  }
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

int pdhg_batch_spectral_norms(pdhg_batch* h, const double* v0, int iterations, double* norms) {
  int rc = check(h, false);
  if (rc) return rc;
  if (!v0 || !norms || iterations < 0) return bfail(h, PDHG_E_ARGUMENT, "null argument");
  BHIP_TRY(h, hipSetDevice(h->device));
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  // The first length in the reference's order (src/la/scaling.cpp:176-178), per LP.
  std::vector<char> trivial(h->nlp, 0);
  for (int32_t k = 0; k < h->nlp; ++k) {
    const LpDesc& d = h->desc_host[k];
    double len = 0.0;
    for (int32_t j = 0; j < d.n; ++j) len += v0[d.col0 + j] * v0[d.col0 + j];
    len = std::sqrt(len);
    const bool empty = d.m == 0 || d.n == 0 || (d.a[0].nitems == 0 && d.a[0].nlr == 0);
    trivial[k] = (empty || len == 0.0) ? 1 : 0;
    h->ctl_host[k].pw_len = len;
    h->ctl_host[k].pw_norm = 0.0;
    h->ctl_host[k].pw_zero = trivial[k];
  }
  BHIP_TRY(h, hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl) * h->nlp, hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(h->sx, v0, sizeof(double) * h->sum_n, hipMemcpyHostToDevice, h->stream));
  launch_power(h, iterations);
  BHIP_TRY(h, hipGetLastError());
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int32_t k = 0; k < h->nlp; ++k) {
    const Ctl& c = h->ctl_host[k];
    // An empty LP, a zero start or a zero vector mid-iteration give 1 (scaling.cpp:166,179).
    norms[k] = (trivial[k] || c.pw_zero || !(c.pw_norm > 0.0)) ? 1.0 : c.pw_norm * 1.01;
  }
  return PDHG_OK;
}

int pdhg_batch_start(pdhg_batch* h, const pdhg_params* p, const double* spectral_norms) {
  int rc = check(h, false);
  if (rc) return rc;
  if (!p || p->iteration_limit < 0) return bfail(h, PDHG_E_ARGUMENT, "bad parameters");
  BHIP_TRY(h, hipSetDevice(h->device));
  if ((rc = start_state(h, p, spectral_norms))) return rc;
  h->chunk = p->chunk_attempts > 0 ? p->chunk_attempts : kDefaultChunkAttempts;
  if ((rc = compact(h, true))) return rc;
  h->captured_running = h->n_act_lp;
  if ((rc = capture(h))) return rc;
  h->started = true;
  return PDHG_OK;
}

int pdhg_batch_tune(pdhg_batch* h, const pdhg_params* p, const double* spectral_norms,
                    int attempts, int32_t* geometry_out) {
  int rc = check(h, false);
  if (rc) return rc;
  if (!p || attempts <= 0) return bfail(h, PDHG_E_ARGUMENT, "bad tuning arguments");
  BHIP_TRY(h, hipSetDevice(h->device));
  const int rounds = std::max(1, (attempts + kSlotAttempts - 1) / kSlotAttempts);
  pdhg_params q = *p;
  q.iteration_limit = std::max<int64_t>(p->iteration_limit, int64_t{rounds} * kSlotAttempts);
  int rate_khz = 0;
  BHIP_TRY(h, hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, h->device));
  if (rate_khz <= 0) return bfail(h, PDHG_E_LAUNCH, "device wall clock rate unavailable");
  // The cooperative floor: the smallest LP alone in the cooperative schedule, device time per
  // attempt (what a cooperative attempt costs once the other LPs have stopped).
  int32_t smallest = 0;
  for (int32_t k = 1; k < h->nlp; ++k) {
    const LpDesc &a = h->desc_host[k], &b = h->desc_host[smallest];
    if (int64_t{a.a[0].nitems} + a.t[0].nitems < int64_t{b.a[0].nitems} + b.t[0].nitems) smallest = k;
  }
  std::fill(h->coop.begin(), h->coop.end(), 0);
  h->coop[smallest] = 1;
  if ((rc = set_bins(h)) || (rc = start_state(h, &q, spectral_norms))) return rc;
  hipEvent_t e0, e1;
  BHIP_TRY(h, hipEventCreate(&e0));
  BHIP_TRY(h, hipEventCreate(&e1));
  BHIP_TRY(h, hipEventRecord(e0, h->stream));
  launch_coop_rounds(h, h->stream, rounds);
  BHIP_TRY(h, hipEventRecord(e1, h->stream));
  BHIP_TRY(h, hipEventSynchronize(e1));
  float ms = 0.0f;
  BHIP_TRY(h, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double coop_us = 1e3 * ms / (rounds * kSlotAttempts);
  // Each persistent geometry: device clock per attempt of each LP.
  std::vector<double> best(h->nlp, std::numeric_limits<double>::infinity());
  std::vector<int32_t> pick(h->nlp, 0);
  std::vector<long long> clk(h->nlp);
  std::fill(h->coop.begin(), h->coop.end(), 0);
  for (int g = 0; g < kGeos; ++g) {
    std::fill(h->geo.begin(), h->geo.end(), g);
    if ((rc = set_bins(h)) || (rc = start_state(h, &q, spectral_norms))) return rc;
    BHIP_TRY(h, hipMemsetAsync(h->clocks, 0, sizeof(long long) * h->nlp, h->stream));
    launch_rounds(h, g, h->stream, rounds);
    BHIP_TRY(h, hipGetLastError());
    BHIP_TRY(h, hipMemcpyAsync(clk.data(), h->clocks, sizeof(long long) * h->nlp, hipMemcpyDeviceToHost, h->stream));
    BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
    BHIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int32_t k = 0; k < h->nlp; ++k) {
      const int64_t done = std::max<int64_t>(h->ctl_host[k].attempts, 1);
      const double us = 1e3 * static_cast<double>(clk[k]) / rate_khz / static_cast<double>(done);
      if (us < best[k]) {
        best[k] = us;
        pick[k] = g;
      }
    }
  }
  // An LP runs alone in a persistent workgroup only when that beats even the cooperative
  // floor; everything else stays cooperative.
  for (int32_t k = 0; k < h->nlp; ++k) {
    h->coop[k] = best[k] < coop_us ? 0 : 1;
    h->geo[k] = pick[k];
    if (geometry_out) geometry_out[k] = h->coop[k] ? -1 : pick[k];
  }
  if ((rc = set_bins(h))) return rc;
  h->started = false;
  if (h->exec) {
    BHIP_TRY(h, hipGraphExecDestroy(h->exec));
    h->exec = nullptr;
  }
  return PDHG_OK;
}

int pdhg_batch_promote(pdhg_batch* h, const pdhg_params* p) {
  int rc = check(h, true);
  if (rc) return rc;
  if (!p || p->iteration_limit < 0) return bfail(h, PDHG_E_ARGUMENT, "bad parameters");
  if (h->precision != PDHG_PRECISION_MIXED || !mixed(h))
    return bfail(h, PDHG_E_STATE, "promote needs a mixed-precision batch still in its FP32 phase");
  BHIP_TRY(h, hipSetDevice(h->device));
  const size_t un = static_cast<size_t>(h->sum_n), um = static_cast<size_t>(h->sum_m);
  if (!h->xb64) {
    if ((rc = balloc(h, &h->xb64, 2 * un)) || (rc = balloc(h, &h->yb64, 2 * um))) return rc;
  }
  const int64_t w = std::max<int64_t>(std::max<int64_t>(2 * h->sum_n, 2 * h->sum_m), 1);
  const int grid = static_cast<int>(std::min<int64_t>((w + 255) / 256, 8192));
  k_batch_widen<<<grid, 256, 0, h->stream>>>(h->xb32, h->xrest, h->xb64, 2 * h->sum_n, h->yb32,
                                            h->yrest, h->yb64, 2 * h->sum_m);
  BHIP_TRY(h, hipGetLastError());
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  for (int32_t k = 0; k < h->nlp; ++k) {
    Ctl& c = h->ctl_host[k];
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
  }
  BHIP_TRY(h, hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(Ctl) * h->nlp, hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  h->arith = PDHG_PRECISION_FP64;
  // Every LP runs again: the cooperative grids take all of them, then the graph is captured
  // with FP64 values.
  if ((rc = compact(h, true))) return rc;
  h->captured_running = h->n_act_lp;
  return capture(h);
}

int pdhg_batch_run(pdhg_batch* h, int chunks, int32_t* running) {
  int rc = check(h, true);
  if (rc) return rc;
  for (int c = 0; c < chunks; ++c) BHIP_TRY(h, hipGraphLaunch(h->exec, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  int32_t r = 0, rc_coop = 0;
  for (int32_t k = 0; k < h->nlp; ++k) {
    r += h->ctl_host[k].halt ? 0 : 1;
    rc_coop += (!h->ctl_host[k].halt && h->coop[k]) ? 1 : 0;
  }
  if (running) *running = r;
  // Shrink the cooperative grids to the LPs still running once a tenth of them have stopped
  // (workgroups of stopped LPs exit at once but still take dispatch slots). Measured on the
  // 512-LP workload: halving 34.5 s of iterations, 0.8 33.9 s, 0.9 33.5 s, 0.95 33.5 s.
  constexpr double kRecompactFraction = 0.9;
  if (rc_coop > 0 && rc_coop <= kRecompactFraction * h->captured_running) {
    if ((rc = compact(h, false))) return rc;
    h->captured_running = rc_coop;
    if ((rc = capture(h))) return rc;
  }
  return PDHG_OK;
}

int pdhg_batch_states(pdhg_batch* h, pdhg_state* out) {
  int rc = check(h, false);
  if (rc) return rc;
  if (!out) return bfail(h, PDHG_E_ARGUMENT, "null output");
  for (int32_t k = 0; k < h->nlp; ++k) {
    const Ctl& c = h->ctl_host[k];
    pdhg_state& s = out[k];
    s.iteration = c.iteration;
    s.attempts = c.attempts;
    s.restarts = c.restarts;
    s.ticks = c.ntick;
    s.done = c.done;
    s.limit_hit = c.limit_hit;
    s.halted = c.halt;
    s.eta = c.eta;
    s.omega = c.omega;
    s.best_worst = c.best_worst;
    s.best_primal = c.best_primal;
    s.best_dual = c.best_dual;
    s.best_dual_objective = c.best_dobj;
  }
  return PDHG_OK;
}

int pdhg_batch_finish(pdhg_batch* h, double* col_value, double* y, double* reduced,
                      double* residuals) {
  int rc = check(h, true);
  if (rc) return rc;
  if (mixed(h)) launch_finish<float>(h, h->sx);
  else launch_finish<double>(h, h->sx);
  BHIP_TRY(h, hipGetLastError());
  if (col_value) BHIP_TRY(h, hipMemcpyAsync(col_value, h->bestx, sizeof(double) * h->sum_n, hipMemcpyDeviceToHost, h->stream));
  if (y) BHIP_TRY(h, hipMemcpyAsync(y, h->besty, sizeof(double) * h->sum_m, hipMemcpyDeviceToHost, h->stream));
  if (reduced) BHIP_TRY(h, hipMemcpyAsync(reduced, h->sx, sizeof(double) * h->sum_n, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Ctl) * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  if (residuals)
    for (int32_t k = 0; k < h->nlp; ++k)
      std::memcpy(residuals + 9 * static_cast<size_t>(k), h->ctl_host[k].final_res, sizeof(double) * 9);
  return PDHG_OK;
}

int pdhg_batch_adjoint(pdhg_batch* h, const double* grad_x, const double* grad_y, double tol,
                       int max_iterations, double* grad_b, double* grad_c, int32_t* iterations,
                       double* residuals) {
  int rc = check(h, true);
  if (rc) return rc;
  if (!grad_x || !grad_y || !grad_b || !grad_c || !(tol > 0.0) || max_iterations < 0)
    return bfail(h, PDHG_E_ARGUMENT, "bad adjoint arguments");
  BHIP_TRY(h, hipSetDevice(h->device));
  const size_t un = static_cast<size_t>(h->sum_n), um = static_cast<size_t>(h->sum_m);
  if (!h->adj_mc) {
    if ((rc = balloc(h, &h->adj_mc, un)) || (rc = balloc(h, &h->adj_mr, um)) ||
        (rc = balloc(h, &h->adj_gx, un)) || (rc = balloc(h, &h->adj_gy, um)) ||
        (rc = balloc(h, &h->adj_gb, um)) || (rc = balloc(h, &h->adj_gc, un)) ||
        (rc = balloc(h, &h->adj_res, 2 * static_cast<size_t>(h->nlp))) ||
        (rc = balloc(h, &h->adj_it, 2 * static_cast<size_t>(h->nlp))))
      return rc;
    for (int k = 0; k < 3; ++k)
      if ((rc = balloc(h, &h->adj_n[k], un)) || (rc = balloc(h, &h->adj_m[k], um))) return rc;
  }
  BHIP_TRY(h, hipMemcpyAsync(h->adj_gx, grad_x, sizeof(double) * un, hipMemcpyHostToDevice, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(h->adj_gy, grad_y, sizeof(double) * um, hipMemcpyHostToDevice, h->stream));
  AdjWork w;
  w.mask_c = h->adj_mc;
  w.mask_r = h->adj_mr;
  w.gx = h->adj_gx;
  w.gy = h->adj_gy;
  w.gb = h->adj_gb;
  w.gc = h->adj_gc;
  w.vn0 = h->adj_n[0];
  w.vn1 = h->adj_n[1];
  w.vn2 = h->adj_n[2];
  w.vm0 = h->adj_m[0];
  w.vm1 = h->adj_m[1];
  w.vm2 = h->adj_m[2];
  w.iters = h->adj_it;
  w.resid = h->adj_res;
  Args<double, double> a = args_of<double, double>(h, 1);
  a.order = h->order_all;
  k_batch_adjoint<Geo1, 1><<<h->nlp, Geo1::kThreads, 0, h->stream>>>(a, w, tol, max_iterations);
  BHIP_TRY(h, hipGetLastError());
  BHIP_TRY(h, hipMemcpyAsync(grad_b, h->adj_gb, sizeof(double) * um, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(grad_c, h->adj_gc, sizeof(double) * un, hipMemcpyDeviceToHost, h->stream));
  if (iterations)
    BHIP_TRY(h, hipMemcpyAsync(iterations, h->adj_it, sizeof(int32_t) * 2 * h->nlp, hipMemcpyDeviceToHost, h->stream));
  if (residuals)
    BHIP_TRY(h, hipMemcpyAsync(residuals, h->adj_res, sizeof(double) * 2 * h->nlp, hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

int pdhg_batch_adjoint_masks(pdhg_batch* h, double* mask_c, double* mask_r) {
  int rc = check(h, true);
  if (rc) return rc;
  if (!mask_c || !mask_r) return bfail(h, PDHG_E_ARGUMENT, "null mask output");
  if (!h->adj_mc) return bfail(h, PDHG_E_STATE, "no adjoint has run on this batch");
  BHIP_TRY(h, hipSetDevice(h->device));
  BHIP_TRY(h, hipMemcpyAsync(mask_c, h->adj_mc, sizeof(double) * static_cast<size_t>(h->sum_n),
                             hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipMemcpyAsync(mask_r, h->adj_mr, sizeof(double) * static_cast<size_t>(h->sum_m),
                             hipMemcpyDeviceToHost, h->stream));
  BHIP_TRY(h, hipStreamSynchronize(h->stream));
  return PDHG_OK;
}

}  // extern "C"
