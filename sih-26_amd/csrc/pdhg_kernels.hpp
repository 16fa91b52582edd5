// pdhg_kernels.hpp - device building blocks of the MI355X restarted-PDHG engine.
//
// One sparse traversal serves every matrix product the engine needs. A matrix is held in
// CSR (rows of Ahat, and separately the rows of Ahat^T, i.e. the columns of Ahat), and a
// host-built plan cuts it into work items, one per 256-thread workgroup:
//
//   stream item  - fewer than kTile nonzeros covering whole rows [r0, r1). The workgroup streams
//                  values and indices with unit stride, multiplies each by the gathered
//                  right-hand side and stages the products in LDS; then each row is summed by
//                  a power-of-two group of lanes (1..64) sized to the rows in the item, and
//                  reduced with cross-lane shuffles. Short rows therefore never leave a lane
//                  idle on a coalesced load.
//   long item    - one segment (< kSeg nonzeros) of a row of kTile or more (the power-law
//                  rows of a Zipf-distributed pattern). The workgroup writes the segment's
//                  sum and nothing else. A single "tail" workgroup launched afterwards adds each
//                  row's segment sums in index order (staged through LDS) and runs that row's
//                  epilogue. Workgroups never wait on or signal each other inside a kernel:
//                  on gfx950 a device-scope release fence writes back the XCD's L2, and one per
//                  workgroup costs more than the product itself. No floating-point atomics
//                  are used anywhere, so results are identical run to run.
//
// The "epilogue" is the per-row work fused after the dot product: the PDHG primal or dual
// update, or a convergence-test term. Each epilogue also feeds per-workgroup partial sums
// that a single-workgroup kernel adds in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sankhya_hip {

// Work-item geometry (measured best on the 200,000 x 500,000 Zipf LP: tiles of 1024 and 4096,
// segments of 2048 to 8192 and 256-row items were all slower).
constexpr int kThreads = 256;   // 4 wave64 wavefronts
constexpr int kTile = 2048;     // nonzeros per stream item
constexpr int kSeg = 4096;      // nonzeros per long-row segment
constexpr int kTail = 1024;     // threads of the single-workgroup tail kernels
constexpr int kChunk = 2048;    // long-row segment sums staged in LDS per tail round
constexpr int kSegBatch = 8;    // long-row segment steps whose gathers are issued together
constexpr int kMaxRows = 512;   // rows per stream item

__device__ __forceinline__ bool finite_bound(double v) {
  // include/sankhya/types.hpp:89-96: absent when infinite or at/after the MPS 1e30 marker.
  return !(__builtin_isinf(v) || fabs(v) >= 1e30);
}

// std::max / std::min argument order: a NaN second argument is ignored, as in the reference.
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ double project(double v, double lo, double hi) {
  if (finite_bound(lo) && v < lo) return lo;
  if (finite_bound(hi) && v > hi) return hi;
  return v;
}

// The matrix is streamed once per product. NT = true loads it non-temporally, so that it does
// not evict the gathered vectors from L2 and the 256 MB MALL (the Infinity Cache); that pays on
// an LP whose matrices do not fit the MALL. When the whole working set fits (NT = false), plain
// loads leave the matrix in the MALL from one product to the next, and the next attempt reads
// it from there: on the 200,000 x 500,000 Zipf LP (about 190 MB on the device) an iteration
// took 58.6 us with plain loads against 62.4 us non-temporal; on the 1,000,000 x 2,000,000 LP
// (about 750 MB) 229 us against 213 us.
template <bool NT, typename T>
__device__ __forceinline__ T mat_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT = true, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  return mat_load<NT>(p);
}
template <bool NT = true>
__device__ __forceinline__ int2 stream_load(const int2* p) {
  const long long v = mat_load<NT>(reinterpret_cast<const long long*>(p));
  return make_int2(static_cast<int>(v & 0xffffffffLL), static_cast<int>(v >> 32));
}
// Two consecutive matrix values (the pointer is pair-aligned), widened to double.
template <bool NT = true>
__device__ __forceinline__ double2 load_vals2(const double* p) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = mat_load<NT>(reinterpret_cast<const d2*>(p));
  return make_double2(v.x, v.y);
}
template <bool NT = true>
__device__ __forceinline__ double2 load_vals2(const float* p) {
  const double v = mat_load<NT>(reinterpret_cast<const double*>(p));
  float2 f;
  __builtin_memcpy(&f, &v, sizeof(f));
  return make_double2(static_cast<double>(f.x), static_cast<double>(f.y));
}

// Cross-lane steps of the butterfly reductions, smallest offset first. After the steps below
// `half`, every aligned group of `half` lanes holds one value, so any lane of the other half of
// the aligned 2*half group is as good a partner as lane ^ half: IEEE addition and smax are
// commutative, so the result is the one an exact xor exchange gives. That lets the steps within
// a row of 16 lanes use DPP lane moves (quad_perm for 1 and 2, row_half_mirror for 4, row_mirror
// for 8), a few cycles each, instead of a round trip through the LDS crossbar.
template <int HALF>
__device__ __forceinline__ double dpp_partner(double v) {
  static_assert(HALF == 1 || HALF == 2 || HALF == 4 || HALF == 8, "DPP steps stay in a row");
  constexpr int kCtrl = HALF == 1 ? 0xB1 : HALF == 2 ? 0x4E : HALF == 4 ? 0x141 : 0x140;
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), kCtrl, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// Sum over each aligned group of L lanes (L a power of two, at most 64; uniform), which every
// lane of the group receives. The lanes of one group are all active or all inactive.
__device__ __forceinline__ double group_sum(double t, int L) {
  if (L >= 2) t += dpp_partner<1>(t);
  if (L >= 4) t += dpp_partner<2>(t);
  if (L >= 8) t += dpp_partner<4>(t);
  if (L >= 16) t += dpp_partner<8>(t);
  if (L >= 32) t += __shfl_xor(t, 16);
  if (L >= 64) t += __shfl_xor(t, 32);
  return t;
}

// Whole-wavefront sum and maximum (every lane active), returned to every lane: DPP within each
// row of 16 lanes, then the four row results read as scalars and combined as the xor-16 and
// xor-32 steps would combine them.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_partner<1>(v);
  v += dpp_partner<2>(v);
  v += dpp_partner<4>(v);
  v += dpp_partner<8>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ __forceinline__ double wave_max(double v) {
  v = smax(v, dpp_partner<1>(v));
  v = smax(v, dpp_partner<2>(v));
  v = smax(v, dpp_partner<4>(v));
  v = smax(v, dpp_partner<8>(v));
  return smax(smax(readlane_d(v, 0), readlane_d(v, 16)), smax(readlane_d(v, 32), readlane_d(v, 48)));
}

// Fixed-shape reduction of NS sums and NM maxima over one workgroup; result in thread 0.
template <int NS, int NM>
struct Acc {
  double s[NS > 0 ? NS : 1];
  double m[NM > 0 ? NM : 1];
  __device__ void zero() {
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
  }
};

// The wave results go through LDS; lane k of wave 0 then combines value k over the waves (in
// wave order), all NS + NM values at once, and thread 0 reads them from those lanes.
template <int NS, int NM>
__device__ void block_reduce(Acc<NS, NM>& a, double* red /* >= waves*(NS+NM) doubles of LDS */) {
  constexpr int NV = NS + NM;
  static_assert(NV <= 64, "one lane per value");
  const int nwaves = static_cast<int>(blockDim.x) >> 6;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NS; ++k) a.s[k] = wave_sum(a.s[k]);
#pragma unroll
  for (int k = 0; k < NM; ++k) a.m[k] = wave_max(a.m[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) red[wave * NV + k] = a.s[k];
#pragma unroll
    for (int k = 0; k < NM; ++k) red[wave * NV + NS + k] = a.m[k];
  }
  __syncthreads();
  if (wave == 0) {
    double t = 0.0;
    if (lane < NV) {
      const bool is_max = lane >= NS;
      t = red[lane];
#pragma unroll 4
      for (int w = 1; w < nwaves; ++w) {
        const double v = red[w * NV + lane];
        t = is_max ? smax(t, v) : t + v;
      }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) a.s[k] = readlane_d(t, k);
#pragma unroll
    for (int k = 0; k < NM; ++k) a.m[k] = readlane_d(t, NS + k);
  }
  __syncthreads();
}

// Sum (or max) of p[t*stride + idx] for t in [0, count), in a fixed order, over one
// workgroup. Every thread returns the total.
__device__ inline double block_sum_strided(const double* p, int64_t count, int stride, int idx,
                                           double* red, bool is_max) {
  const int nwaves = static_cast<int>(blockDim.x) >> 6;
  double t = 0.0;
  for (int64_t q = threadIdx.x; q < count; q += blockDim.x) {
    const double v = p[q * stride + idx];
    t = is_max ? smax(t, v) : t + v;
  }
  t = is_max ? wave_max(t) : wave_sum(t);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < nwaves; ++w) r = is_max ? smax(r, red[w]) : r + red[w];
  __syncthreads();
  return r;
}

// NV sums (or maxima, where is_max[k]) over the concatenation of a[0..ca) and b[0..cb), each
// entry NV wide at the given stride, in one pass: every lane walks its share in a fixed order,
// then one cross-lane and one cross-wave step. Every thread returns the totals.
template <int NV>
__device__ inline void block_sums(const double* a, int64_t ca, const double* b, int64_t cb,
                                  int stride, const bool (&is_max)[NV], double (&out)[NV],
                                  double* red /* >= waves*NV */) {
  const int nwaves = static_cast<int>(blockDim.x) >> 6;
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = 0.0;
  const int64_t total = ca + cb;
  constexpr int U = 4;  // entries per lane whose loads are issued before any is used
  for (int64_t q0 = threadIdx.x; q0 < total; q0 += static_cast<int64_t>(blockDim.x) * U) {
    double v[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = q0 + static_cast<int64_t>(u) * blockDim.x;
      const double* row = q < ca ? a + q * stride : b + (q - ca) * stride;
#pragma unroll
      for (int k = 0; k < NV; ++k) v[u][k] = q < total ? row[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < NV; ++k) t[k] = is_max[k] ? smax(t[k], v[u][k]) : t[k] + v[u][k];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = is_max[k] ? wave_max(t[k]) : wave_sum(t[k]);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[(threadIdx.x >> 6) * NV + k] = t[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double r = red[k];
    for (int w = 1; w < nwaves; ++w) r = is_max[k] ? smax(r, red[w * NV + k]) : r + red[w * NV + k];
    out[k] = r;
  }
  __syncthreads();
}

// NV sums (or maxima, where is_max[k]) of two arrays at once, each holding rounds * kSumU *
// blockDim entries NV wide at the given stride, zero past the items: every load unconditional,
// the loads of both arrays issued together (one memory round trip per round), then the
// reductions. Per lane the entries are taken in index order, as block_sums takes them. The
// totals are returned in wave 0 (the caller reads them in thread 0).
template <int NV>
__device__ inline void block_sums2_padded(const double* a, const double* b, int64_t rounds,
                                          int stride, const bool (&is_max)[NV],
                                          double (&out_a)[NV], double (&out_b)[NV],
                                          double* red /* >= waves*NV */) {
  constexpr int U = 4;
  const int nwaves = static_cast<int>(blockDim.x) >> 6;
  double ta[NV], tb[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) ta[k] = tb[k] = 0.0;
  for (int64_t r = 0; r < rounds; ++r) {
    double va[U][NV], vb[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = (r * U + u) * static_cast<int64_t>(blockDim.x) + threadIdx.x;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        va[u][k] = a[q * stride + k];
        vb[u][k] = b[q * stride + k];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        ta[k] = is_max[k] ? smax(ta[k], va[u][k]) : ta[k] + va[u][k];
        tb[k] = is_max[k] ? smax(tb[k], vb[u][k]) : tb[k] + vb[u][k];
      }
    }
  }
  // Wave results through LDS; lane k of wave 0 combines value k over the waves in wave order,
  // and the totals are read from those lanes: they are valid in wave 0 (thread 0 uses them).
  const int lane = threadIdx.x & 63;
  auto finish = [&](double (&t)[NV], double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) t[k] = is_max[k] ? wave_max(t[k]) : wave_sum(t[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) red[(threadIdx.x >> 6) * NV + k] = t[k];
    }
    __syncthreads();
    if ((threadIdx.x >> 6) == 0) {
      double v = 0.0;
      if (lane < NV) {
        bool mx = false;
#pragma unroll
        for (int k = 0; k < NV; ++k) mx = (lane == k) ? is_max[k] : mx;
        v = red[lane];
#pragma unroll 4
        for (int w = 1; w < nwaves; ++w) v = mx ? smax(v, red[w * NV + lane]) : v + red[w * NV + lane];
      }
#pragma unroll
      for (int k = 0; k < NV; ++k) out[k] = readlane_d(v, k);
    }
    __syncthreads();
  };
  finish(ta, out_a);
  finish(tb, out_b);
}

// A CSR matrix with its work-item plan, as kernels see it.
template <typename V>
struct PlanView {
  const int32_t* rowptr;
  const int32_t* col;
  const V* val;
  const int4* items;   // {r0, r1, k0, k1}; r1 < 0: a long-row segment {seg id, -lr-1, k0, k1}
  const int4* lr_info;  // per long row {row, first segment, segment count, 0}
  const int4* lr_groups;  // {first, last) long rows, first segment, segment count} per LDS round
  double* seg_part;      // segment sums, NR per segment
  int32_t nitems;
  int32_t nlr;
  int32_t ngroups;
  // lr_groups[0] by value when the host knows it (has_group0 = 1), so the tail kernels need
  // not load it before their first segment sums: one memory round trip less per launch.
  int4 group0;
  int32_t has_group0;
  // Staged epilogue inputs, kStage doubles per long row, or null. When set, the workgroup of a
  // long row's first segment loads that row's epilogue inputs beside its own work and stores
  // them here, and the tail kernel reads them by long-row index: the tail then needs no load
  // that waits on another (the row id from lr_info, then the row's inputs), one memory round
  // trip less on every attempt's critical path. Valid only where the epilogue's inputs do not
  // change between the product kernel and its tail kernel.
  double* lr_in = nullptr;
};

constexpr int kStage = 8;  // doubles staged per long row (every epilogue's In fits)

// An epilogue input record kept as doubles.
template <class In>
constexpr bool kStageable = sizeof(In) >= sizeof(double) && sizeof(In) % sizeof(double) == 0 &&
                            sizeof(In) <= kStage * sizeof(double);
template <class In>
__device__ __forceinline__ void stage_store(double* p, const In& in) {
  double v[sizeof(In) / sizeof(double)];
  __builtin_memcpy(v, &in, sizeof(In));
#pragma unroll
  for (int k = 0; k < static_cast<int>(sizeof(In) / sizeof(double)); ++k) p[k] = v[k];
}
template <class In>
__device__ __forceinline__ In stage_load(const double* p) {
  double v[sizeof(In) / sizeof(double)];
#pragma unroll
  for (int k = 0; k < static_cast<int>(sizeof(In) / sizeof(double)); ++k) v[k] = p[k];
  In r;
  __builtin_memcpy(&r, v, sizeof(In));
  return r;
}

template <int NR>
struct RowSum {
  double v[NR];
};

// Work-item geometry of one kernel family: threads per workgroup, nonzeros per stream item and
// per long-row segment, rows per stream item. kRowsPerLane is how many rows one lane finishes
// in a stream item at most: a lane group of L lanes owns rows grp, grp + ngrp, ...; with L > 1
// an item has at most ngrp rows, with L = 1 at most kMaxRows.
template <int NT, int TILE, int SEG, int MAXROWS>
struct Geometry {
  static constexpr int kThreads = NT, kTile = TILE, kSeg = SEG, kMaxRows = MAXROWS;
  static constexpr int kRowsPerLane = MAXROWS / NT;
  static_assert(MAXROWS % NT == 0 && TILE % (2 * NT) == 0, "geometry");
};
// The single-LP engine: 256-thread workgroups, one item each.
using SingleGeo = Geometry<kThreads, kTile, kSeg, kMaxRows>;

// Process one item. gather(j) returns RowSum<NR> for column j. The epilogue type E provides
// E::In load(row) (the row's own inputs: iterate, bounds, running sum) and
// operator()(row, sum, in, acc), run once per completed row in one lane. A long-row segment
// only stores its sum; finish_long_rows completes those rows later.
//
// Memory round trips: the item descriptor arrives with the caller's control-block loads; then
// ONE round trip brings the item's nonzeros, its row extents and every epilogue input; then the
// gathers. Products are staged in LDS and each row is summed by L lanes with shuffles.
template <class GEO, typename V, int NR, bool NT = true, int NS, int NM, class G, class E>
__device__ void run_item_geo(const PlanView<V>& P, const int4 it, G& gather, E& epi,
                             Acc<NS, NM>& acc, double* lds, double* red) {
  constexpr int kThreads = GEO::kThreads, kTile = GEO::kTile, kSeg = GEO::kSeg;
  constexpr int kRowsPerLane = GEO::kRowsPerLane;
  const int tid = threadIdx.x;
  if (it.y >= 0) {
    const int r0 = it.x, r1 = it.y, k0 = it.z, k1 = it.w;
    const int rows = r1 - r0;
    int L = 64;
    while (L > 1 && rows * L > kThreads) L >>= 1;
    const int grp = tid / L, lane = tid & (L - 1), ngrp = kThreads / L;
    // Two nonzeros per lane per step (one 16-byte value load, one 8-byte index load), from a
    // pair-aligned base; kTile / (2 kThreads) steps, all loads issued before the first gather.
    // The plan keeps an item below kTile nonzeros, so its span from the aligned base fits.
    const int kb = k0 & ~1;
    constexpr int U = kTile / (2 * kThreads);
    int2 cidx[U];
    double2 aval[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + 2 * (u * kThreads + tid);
      if (k < k1) {
        cidx[u] = stream_load<NT>(reinterpret_cast<const int2*>(P.col + k));
        aval[u] = load_vals2<NT>(P.val + k);
      }
    }
    int rs[kRowsPerLane], re[kRowsPerLane];
    typename E::In ein[kRowsPerLane];
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
      rs[q] = 0;
      re[q] = 0;
      const int rr = grp + q * ngrp;
      if (lane == 0 && rr < rows) {
        rs[q] = P.rowptr[r0 + rr];
        re[q] = P.rowptr[r0 + rr + 1];
        ein[q] = epi.load(r0 + rr);
      }
    }
    // All gathers go out before any product is formed: a lane whose nonzero lies outside the
    // item gathers column 0 (present in every matrix that has an item) and never uses it, so
    // no branch surrounds a load and the gathers of all U steps are in flight together.
    RowSum<NR> g0[U], g1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + 2 * (u * kThreads + tid);
      g0[u] = gather((k < k1 && k >= k0) ? cidx[u].x : 0);
      g1[u] = gather((k + 1 < k1) ? cidx[u].y : 0);
    }
    // Products of right-hand side 0 go to LDS; those of the others wait in registers and
    // reuse the same kTile doubles of LDS one after another (NR passes, one buffer).
    double hold[NR > 1 ? NR - 1 : 1][U][2];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + 2 * (u * kThreads + tid);
      if (k < k1) {
        if (k >= k0) {
          lds[k - k0] = aval[u].x * g0[u].v[0];
#pragma unroll
          for (int r = 1; r < NR; ++r) hold[r - 1][u][0] = aval[u].x * g0[u].v[r];
        }
        if (k + 1 < k1) {
          lds[k + 1 - k0] = aval[u].y * g1[u].v[0];
#pragma unroll
          for (int r = 1; r < NR; ++r) hold[r - 1][u][1] = aval[u].y * g1[u].v[r];
        }
      }
    }
    const int src = (tid & 63) & ~(L - 1);  // the group's first lane within the wavefront
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
      if (L > 1) {
        rs[q] = __shfl(rs[q], src);
        re[q] = __shfl(re[q], src);
      }
      rs[q] -= k0;
      re[q] -= k0;
    }
    RowSum<NR> sum[kRowsPerLane];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r > 0) {
        __syncthreads();  // every lane has read pass r-1
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = kb + 2 * (u * kThreads + tid);
          if (k < k1) {
            if (k >= k0) lds[k - k0] = hold[r - 1][u][0];
            if (k + 1 < k1) lds[k + 1 - k0] = hold[r - 1][u][1];
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kRowsPerLane; ++q) {
        const int rr = grp + q * ngrp;
        double t = 0.0;
        if (rr < rows) {
          for (int s = rs[q] + lane; s < re[q]; s += L) t += lds[s];
          t = group_sum(t, L);
        }
        sum[q].v[r] = t;
      }
    }
#pragma unroll
    for (int q = 0; q < kRowsPerLane; ++q) {
      const int rr = grp + q * ngrp;
      if (rr < rows && lane == 0) epi(r0 + rr, sum[q], ein[q], acc);
    }
    return;
  }
  // Long-row segment {global segment id, -(long row id)-1, k0, k1}: kSeg / kThreads
  // independent loads per lane (a segment is below kSeg nonzeros, so its span from the aligned
  // base fits kSeg), then the gathers.
  const int seg_id = it.x, k0 = it.z, k1 = it.w;
  Acc<NR, 0> seg;
#pragma unroll
  for (int r = 0; r < NR; ++r) seg.s[r] = 0.0;
  // Staging (PlanView::lr_in): the long row's descriptor is read by every segment workgroup
  // (one uniform load beside the nonzeros, unconditional, so nothing waits for it early); the
  // first segment's workgroup loads the row's epilogue inputs once its products are formed and
  // stores them after the reduction.
  constexpr bool kStaged = kStageable<typename E::In>;
  const int lr = -it.y - 1;
  const int4 info = P.lr_info[lr];
  {
    const int kb = k0 & ~1;
    constexpr int U = kSeg / (2 * kThreads);
    int2 cidx[U];
    double2 aval[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kb + 2 * (u * kThreads + tid);
      if (k < k1) {
        cidx[u] = stream_load<NT>(reinterpret_cast<const int2*>(P.col + k));
        aval[u] = load_vals2<NT>(P.val + k);
      }
    }
    // Gathers in batches of kBatch steps, each batch issued whole before its products are
    // added (out-of-segment lanes gather column 0 and add nothing), in the original order.
    constexpr int kBatch = kSegBatch;
#pragma unroll
    for (int ub = 0; ub < U; ub += kBatch) {
      RowSum<NR> g0[kBatch], g1[kBatch];
#pragma unroll
      for (int v = 0; v < kBatch; ++v) {
        const int u = ub + v;
        if (u < U) {
          const int k = kb + 2 * (u * kThreads + tid);
          g0[v] = gather((k < k1 && k >= k0) ? cidx[u].x : 0);
          g1[v] = gather((k + 1 < k1) ? cidx[u].y : 0);
        }
      }
#pragma unroll
      for (int v = 0; v < kBatch; ++v) {
        const int u = ub + v;
        if (u >= U) continue;
        const int k = kb + 2 * (u * kThreads + tid);
        if (k < k1) {
          if (k >= k0) {
#pragma unroll
            for (int r = 0; r < NR; ++r) seg.s[r] += aval[u].x * g0[v].v[r];
          }
          if (k + 1 < k1) {
#pragma unroll
            for (int r = 0; r < NR; ++r) seg.s[r] += aval[u].y * g1[v].v[r];
          }
        }
      }
    }
  }
  const bool stager = kStaged && P.lr_in != nullptr && tid == 0 && seg_id == info.y;
  typename E::In staged{};
  if constexpr (kStaged) {
    if (stager) staged = epi.load(info.x);
  }
  block_reduce(seg, red);
  if (tid == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) P.seg_part[(int64_t)seg_id * NR + r] = seg.s[r];
  }
  if constexpr (kStaged) {
    if (stager) stage_store(P.lr_in + static_cast<int64_t>(lr) * kStage, staged);
  }
}

template <typename V, int NR, bool NT = true, int NS, int NM, class G, class E>
__device__ __forceinline__ void run_item(const PlanView<V>& P, const int4 it, G& gather, E& epi,
                                         Acc<NS, NM>& acc, double* lds, double* red) {
  run_item_geo<SingleGeo, V, NR, NT>(P, it, gather, epi, acc, lds, red);
}

// Completes every long row of P in one workgroup (launched after the kernel that wrote the
// segment sums): for each group of at most blockDim long rows, stage their segment sums in
// LDS with coalesced loads while each row's lane fetches the row's epilogue inputs (from
// P.lr_in when the product kernel staged them); then a power-of-two group of lanes per row adds
// its segments in index order and one lane runs epi.
// `pre()` runs once, after the first group's loads are issued and before they are used, so a
// caller's own independent loads share that memory round trip. It returns false to end the
// walk before anything is written (the same answer in every thread); finish_long_rows_geo then
// returns false.
struct NoPre {
  __device__ bool operator()() const { return true; }
};

// One group of long rows {first, last) long rows, first segment, segment count}: the segment
// sums, row descriptors and epilogue inputs are all loaded, then `hook()`, then LDS, sums and
// epilogues. Returns false (having written nothing) when hook() does. With kFromStage the
// epilogue inputs come from P.lr_in (the product kernel staged them) and every load is
// unconditional at a clamped index, so no load waits for another and nothing waits before all
// of them are issued.
template <int NTAIL, typename V, int NR, bool kFromStage, int NS, int NM, class E, class Hook>
__device__ __forceinline__ bool long_row_group(const PlanView<V>& P, const int4 grp, E& epi,
                                               Acc<NS, NM>& acc, double* lds, Hook& hook) {
  static_assert(!kFromStage || kStageable<typename E::In>, "staged inputs need a stageable In");
  const int nrows = grp.y - grp.x, base0 = grp.z, cnt = grp.w;
  int L = 64;
  while (L > 1 && nrows * L > static_cast<int>(blockDim.x)) L >>= 1;
  const int gi = threadIdx.x / L, lane = threadIdx.x & (L - 1);
  const bool mine = gi < nrows;  // at most one row per lane group: nrows <= blockDim / L
  const int lr = grp.x + (mine ? gi : 0);
  // kChunk * NR / NTAIL loads per lane at most, held in registers until `hook` has run.
  constexpr int U = (kChunk * NR + NTAIL - 1) / NTAIL;
  double v[U];
  int4 info;
  typename E::In ein;
  if constexpr (kFromStage) {
    info = P.lr_info[lr];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
      const double x = P.seg_part[(int64_t)base0 * NR + (t < cnt * NR ? t : 0)];
      v[u] = t < cnt * NR ? x : 0.0;
    }
    ein = stage_load<typename E::In>(P.lr_in + static_cast<int64_t>(lr) * kStage);
  } else {
    info = make_int4(0, 0, 0, 0);
    if (mine) info = P.lr_info[lr];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
      v[u] = t < cnt * NR ? P.seg_part[(int64_t)base0 * NR + t] : 0.0;
    }
    if (mine && lane == 0) ein = epi.load(info.x);
  }
  if (!hook()) return false;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int t = u * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (t < cnt * NR) lds[t] = v[u];
  }
  __syncthreads();
  if (mine) {
    const int s0 = info.y - base0, ns = info.z;
    RowSum<NR> sum;
#pragma unroll
    for (int r = 0; r < NR; ++r) sum.v[r] = 0.0;
#pragma unroll 4
    for (int s = lane; s < ns; s += L) {
#pragma unroll
      for (int r = 0; r < NR; ++r) sum.v[r] += lds[(s0 + s) * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) sum.v[r] = group_sum(sum.v[r], L);
    if (lane == 0) epi(info.x, sum, ein, acc);
  }
  __syncthreads();
  return true;
}

// kFromStage: the single-LP engine's iteration tails, whose product kernels staged the long
// rows' epilogue inputs (PlanView::lr_in) and whose plans carry their first group by value.
// The first group is walked outside the loop, so its loads are issued at kernel start.
template <int NTAIL, typename V, int NR, bool kFromStage = false, int NS, int NM, class E,
          class Pre = NoPre>
__device__ bool finish_long_rows_geo(const PlanView<V>& P, E& epi, Acc<NS, NM>& acc,
                                     double* lds /* kChunk * NR */, Pre pre = Pre()) {
  if (P.ngroups <= 0) return pre();
  const int4 g0 = (kFromStage || P.has_group0) ? P.group0 : P.lr_groups[0];
  if (!long_row_group<NTAIL, V, NR, kFromStage>(P, g0, epi, acc, lds, pre)) return false;
  NoPre none;
  for (int g = 1; g < P.ngroups; ++g)
    long_row_group<NTAIL, V, NR, kFromStage>(P, P.lr_groups[g], epi, acc, lds, none);
  return true;
}

// The single-LP engine's tail kernels run kTail threads.
template <typename V, int NR, int NS, int NM, class E, class Pre = NoPre>
__device__ __forceinline__ bool finish_long_rows(const PlanView<V>& P, E& epi, Acc<NS, NM>& acc,
                                                 double* lds, Pre pre = Pre()) {
  return finish_long_rows_geo<kTail, V, NR>(P, epi, acc, lds, pre);
}
// The iteration tails of the single-LP engine: inputs staged by the product kernel.
template <typename V, int NR, int NS, int NM, class E, class Pre>
__device__ __forceinline__ bool finish_staged_rows(const PlanView<V>& P, E& epi, Acc<NS, NM>& acc,
                                                   double* lds, Pre pre) {
  return finish_long_rows_geo<kTail, V, NR, true>(P, epi, acc, lds, pre);
}

// Adds p[q * stride + k] for q in [0, count) into out[k], k < NV, with the loads of several
// entries per lane issued before they are used. Every lane keeps its own partial sums.
template <int NV>
__device__ __forceinline__ void lane_sums(const double* p, int64_t count, int stride,
                                          double (&out)[NV]) {
  constexpr int U = 4;
  for (int64_t q0 = threadIdx.x; q0 < count; q0 += static_cast<int64_t>(blockDim.x) * U) {
    double v[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = q0 + static_cast<int64_t>(u) * blockDim.x;
#pragma unroll
      for (int k = 0; k < NV; ++k) v[u][k] = q < count ? p[q * stride + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int k = 0; k < NV; ++k) out[k] += v[u][k];
    }
  }
}

// lane_sums over two arrays at once, one value per entry of a (count ca) and two per entry of b
// (count cb): every load of a round is issued before any is used, so both arrays share each
// memory round trip. Every lane keeps its own partial sums.
__device__ __forceinline__ void lane_sums_1_2(const double* a, int64_t ca, const double* b,
                                              int64_t cb, double (&sa)[1], double (&sb)[2]) {
  constexpr int U = 4;
  const int64_t most = ca > cb ? ca : cb;
  const int64_t rounds = (most + static_cast<int64_t>(blockDim.x) * U - 1) /
                         (static_cast<int64_t>(blockDim.x) * U);
  for (int64_t r = 0; r < rounds; ++r) {
    double va[U], vb[U][2];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = (r * U + u) * static_cast<int64_t>(blockDim.x) + threadIdx.x;
      va[u] = q < ca ? a[q] : 0.0;
      vb[u][0] = q < cb ? b[2 * q] : 0.0;
      vb[u][1] = q < cb ? b[2 * q + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sa[0] += va[u];
      sb[0] += vb[u][0];
      sb[1] += vb[u][1];
    }
  }
}

// lane_sums_1_2 over arrays padded with zeros to whole rounds of kSumU * blockDim entries: a
// holds `rounds` such rounds and b twice that, rounds >= 1. Every load is unconditional, and
// the first round is issued before any loop, beside whatever the caller issued before: no load
// waits for another. Same order of additions as lane_sums_1_2, so the same sums bit for bit.
constexpr int kSumU = 4;
__device__ __forceinline__ void lane_sums_1_2_padded(const double* a, const double* b,
                                                     int64_t rounds, double (&sa)[1],
                                                     double (&sb)[2]) {
  auto round = [&](int64_t r) {
    double va[kSumU];
    double2 vb[kSumU];
#pragma unroll
    for (int u = 0; u < kSumU; ++u) {
      const int64_t q = (r * kSumU + u) * static_cast<int64_t>(blockDim.x) + threadIdx.x;
      va[u] = a[q];
      vb[u] = reinterpret_cast<const double2*>(b)[q];
    }
#pragma unroll
    for (int u = 0; u < kSumU; ++u) {
      sa[0] += va[u];
      sb[0] += vb[u].x;
      sb[1] += vb[u].y;
    }
  };
  round(0);
  for (int64_t r = 1; r < rounds; ++r) round(r);
}

}  // namespace sankhya_hip
