// pdhg_kernels.hpp - device building blocks of the MI355X restarted-PDHG engine.
//
// One sparse traversal serves every matrix product the engine needs. A matrix is held in
// CSR (rows of Ahat, and separately the rows of Ahat^T, i.e. the columns of Ahat), and a
// host-built plan cuts it into work items, one per 256-thread workgroup:
//
//   stream item  - up to kTile nonzeros covering whole rows [r0, r1). The workgroup streams
//                  values and indices with unit stride, multiplies each by the gathered
//                  right-hand side and stages the products in LDS; then each row is summed by
//                  a power-of-two group of lanes (1..64) sized to the rows in the item, and
//                  reduced with cross-lane shuffles. Short rows therefore never leave a lane
//                  idle on a coalesced load.
//   long item    - one segment of kSeg nonzeros of a row longer than kTile (the power-law
//                  rows of a Zipf-distributed pattern). Each workgroup writes its segment sum;
//                  the last of the row's workgroups to finish (an integer counter decides which,
//                  not the order of the sum) adds the segment sums in index order and runs the
//                  row's epilogue. No floating-point atomics are used anywhere, so results are
//                  identical run to run.
//
// The "epilogue" is the per-row work fused after the dot product: the PDHG primal or dual
// update, or a convergence-test term. Each epilogue also feeds per-workgroup partial sums
// that a single-workgroup kernel adds in index order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sankhya_hip {

constexpr int kThreads = 256;   // 4 wave64 wavefronts
constexpr int kTile = 2048;     // nonzeros per stream item
constexpr int kSeg = 8192;      // nonzeros per long-row segment
constexpr int kMaxRows = 256;   // rows per stream item

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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = smax(v, __shfl_xor(v, off));
  return v;
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

template <int NS, int NM>
__device__ void block_reduce(Acc<NS, NM>& a, double* red /* >= 4*(NS+NM) doubles of LDS */) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NS; ++k) a.s[k] = wave_sum(a.s[k]);
#pragma unroll
  for (int k = 0; k < NM; ++k) a.m[k] = wave_max(a.m[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) red[wave * (NS + NM) + k] = a.s[k];
#pragma unroll
    for (int k = 0; k < NM; ++k) red[wave * (NS + NM) + NS + k] = a.m[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double t = red[k];
      for (int w = 1; w < kThreads / 64; ++w) t += red[w * (NS + NM) + k];
      a.s[k] = t;
    }
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      double t = red[NS + k];
      for (int w = 1; w < kThreads / 64; ++w) t = smax(t, red[w * (NS + NM) + NS + k]);
      a.m[k] = t;
    }
  }
  __syncthreads();
}

// Sum (or max) of p[t*stride + idx] for t in [0, count), in a fixed order, over one
// workgroup. Every thread returns the total.
__device__ inline double block_sum_strided(const double* p, int64_t count, int stride, int idx,
                                           double* red, bool is_max) {
  double t = 0.0;
  for (int64_t q = threadIdx.x; q < count; q += kThreads) {
    const double v = p[q * stride + idx];
    t = is_max ? smax(t, v) : t + v;
  }
  t = is_max ? wave_max(t) : wave_sum(t);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < kThreads / 64; ++w) r = is_max ? smax(r, red[w]) : r + red[w];
  __syncthreads();
  return r;
}

// A CSR matrix with its work-item plan, as kernels see it.
template <typename V>
struct PlanView {
  const int32_t* rowptr;
  const int32_t* col;
  const V* val;
  const int4* items;  // {r0, r1, k0, k1}; r1 < 0 marks a long-row segment of row r0 with id -r1-1
  const int32_t* lr_nseg;
  const int32_t* lr_base;
  double* seg_part;   // segment sums, NR per segment
  int32_t* lr_count;  // arrival counters, one per long row, zero between launches
  int32_t nitems;
  int32_t nlr;
};

template <int NR>
struct RowSum {
  double v[NR];
};

// Process this workgroup's item. gather(j) returns RowSum<NR> for column j; epi(row, sum,
// acc) runs once per completed row (in one thread). Long-row epilogues accumulate into
// `lr_acc`, which the caller stores per long row so its contribution does not depend on which
// workgroup finished the row. Returns the long-row id whose epilogue ran here, or -1.
template <typename V, int NR, int NS, int NM, class G, class E>
__device__ int run_item(const PlanView<V>& P, G& gather, E& epi, Acc<NS, NM>& acc,
                        Acc<NS, NM>& lr_acc, double* lds, double* red) {
  const int4 it = P.items[blockIdx.x];
  const int tid = threadIdx.x;
  if (it.y >= 0) {
    const int r0 = it.x, r1 = it.y, k0 = it.z, k1 = it.w;
    const int cnt = k1 - k0;
    for (int t = tid; t < cnt; t += kThreads) {
      const int k = k0 + t;
      const double a = static_cast<double>(P.val[k]);
      const RowSum<NR> g = gather(P.col[k]);
#pragma unroll
      for (int r = 0; r < NR; ++r) lds[r * kTile + t] = a * g.v[r];
    }
    __syncthreads();
    const int rows = r1 - r0;
    int L = 64;
    while (L > 1 && rows * L > kThreads) L >>= 1;
    const int grp = tid / L, lane = tid & (L - 1), ngrp = kThreads / L;
    for (int rr = grp; rr < rows; rr += ngrp) {
      const int row = r0 + rr;
      const int s0 = P.rowptr[row] - k0, s1 = P.rowptr[row + 1] - k0;
      RowSum<NR> sum;
#pragma unroll
      for (int r = 0; r < NR; ++r) sum.v[r] = 0.0;
      for (int s = s0 + lane; s < s1; s += L) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sum.v[r] += lds[r * kTile + s];
      }
      for (int off = L >> 1; off > 0; off >>= 1) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sum.v[r] += __shfl_xor(sum.v[r], off);
      }
      if (lane == 0) epi(row, sum, acc);
    }
    return -1;
  }
  // Long-row segment.
  const int row = it.x, lr = -it.y - 1, k0 = it.z, k1 = it.w;
  Acc<NR, 0> seg;
#pragma unroll
  for (int r = 0; r < NR; ++r) seg.s[r] = 0.0;
  for (int k = k0 + tid; k < k1; k += kThreads) {
    const double a = static_cast<double>(P.val[k]);
    const RowSum<NR> g = gather(P.col[k]);
#pragma unroll
    for (int r = 0; r < NR; ++r) seg.s[r] += a * g.v[r];
  }
  block_reduce(seg, red);
  const int nseg = P.lr_nseg[lr];
  const int base = P.lr_base[lr];
  const int s = (k0 - P.rowptr[row]) / kSeg;
  __shared__ int last;
  if (tid == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) P.seg_part[(int64_t)(base + s) * NR + r] = seg.s[r];
    __threadfence();
    const int old = atomicAdd(&P.lr_count[lr], 1);
    last = (old == nseg - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!last) return -1;
  __threadfence();
  RowSum<NR> total;
#pragma unroll
  for (int r = 0; r < NR; ++r)
    total.v[r] = block_sum_strided(P.seg_part + (int64_t)base * NR, nseg, NR, r, red, false);
  if (tid == 0) {
    lr_acc.zero();
    epi(row, total, lr_acc);
    atomicExch(&P.lr_count[lr], 0);
  }
  return lr;
}

}  // namespace sankhya_hip
