// scaling_waves.hpp - device pieces of the preconditioner (src/la/scaling.cpp:14-159) shared by
// the single-LP engine (pdhg_device.hip) and the batched engine (pdhg_batch.hip): one wave per
// matrix line for the Ruiz maxima, the scaling of the entries and the Pock-Chambolle sums in
// storage order, and the flat per-coordinate kernels (factors, fills, scaled vectors). The
// results equal the reference's host passes bit for bit: maxima do not depend on order, the
// sums are added in storage order by one lane, and factors and scaled values use the same
// correctly rounded IEEE double operations in the same order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdhg_kernels.hpp"

namespace sankhya_hip {

// The kernels walk a matrix's work plan (the same items the products use): a stream item's
// rows one wave each, a long row's segments (at most kSeg entries) one workgroup each, so a
// dense row of hundreds of thousands of entries is spread over many workgroups. kScU coalesced
// loads per lane are in flight per step.
constexpr int kScU = 8;

// Largest |a| of entries [k0, k1), one wave; every lane returns it.
__device__ __forceinline__ double wave_absmax(const double* val, int32_t k0, int32_t k1, int lane) {
  double r = 0.0;
  for (int32_t base = k0; base < k1; base += 64 * kScU) {
    double v[kScU];
#pragma unroll
    for (int u = 0; u < kScU; ++u) {
      const int32_t k = base + u * 64 + lane;
      v[u] = k < k1 ? val[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kScU; ++u) r = smax(r, fabs(v[u]));
  }
  for (int off = 32; off > 0; off >>= 1) r = smax(r, __shfl_xor(r, off));
  return r;
}

// Entries [k0, k1) become (a * r_i) * c_j, one wave.
__device__ __forceinline__ void wave_apply(const int32_t* idx, double* val, int32_t k0, int32_t k1,
                                           int lane, double lf, const double* idx_pass,
                                           int line_is_row) {
  for (int32_t base = k0; base < k1; base += 64 * kScU) {
    int32_t id[kScU];
    double v[kScU], other[kScU];
#pragma unroll
    for (int u = 0; u < kScU; ++u) {
      const int32_t k = base + u * 64 + lane;
      id[u] = k < k1 ? idx[k] : 0;  // index 0 exists: this line has an entry
      v[u] = k < k1 ? val[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kScU; ++u) other[u] = idx_pass[id[u]];
#pragma unroll
    for (int u = 0; u < kScU; ++u) {
      const int32_t k = base + u * 64 + lane;
      const double r = line_is_row ? lf : other[u], c = line_is_row ? other[u] : lf;
      if (k < k1) val[k] = v[u] * r * c;
    }
  }
}

// Sum of |a| over [k0, k1) in storage order, one wave: the wave stages a step of entries in
// its LDS slice with coalesced loads, then lane 0 adds them one by one in order, so the result
// is the sequential sum bit for bit. Only lane 0's result is meaningful.
__device__ __forceinline__ double wave_abssum_ordered(const double* val, int32_t k0, int32_t k1,
                                                      int lane, double* lds /* 64 kScU */) {
  double r = 0.0;
  for (int32_t base = k0; base < k1; base += 64 * kScU) {
    double v[kScU];
#pragma unroll
    for (int u = 0; u < kScU; ++u) {
      const int32_t k = base + u * 64 + lane;
      v[u] = k < k1 ? fabs(val[k]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kScU; ++u) lds[u * 64 + lane] = v[u];
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      const int32_t count = k1 - base < 64 * kScU ? k1 - base : 64 * kScU;
      int32_t e = 0;
      for (; e + 8 <= count; e += 8) {  // eight reads in flight, then eight adds in order
        double t[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) t[q] = lds[e + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) r += t[q];
      }
      for (; e < count; ++e) r += lds[e];
    }
    __builtin_amdgcn_wave_barrier();
  }
  return r;
}

namespace {

// This pass's factor 1/sqrt(stat) (1 for an empty line) and the running product d *= factor.
__global__ __launch_bounds__(kThreads) void k_sc_factor(const double* stat, double* pass, double* d,
                                                        int32_t count) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const double s = stat[i];
    const double f = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
    pass[i] = f;
    d[i] *= f;
  }
}

__global__ __launch_bounds__(kThreads) void k_sc_fill(double* v, int32_t count, double value) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    v[i] = value;
}

// Scaled cost and bounds: c^ = c Dc, l^ = l (1/Dc), u^ = u (1/Dc), rl^ = rl Dr, ru^ = ru Dr;
// infinite sides stay infinite (scaling.cpp:105-107,139-157).
__global__ __launch_bounds__(kThreads) void k_sc_vectors(int32_t n, int32_t m, const double* dc,
                                                         const double* dr, const double* cost,
                                                         const double* cl, const double* cu,
                                                         const double* rl, const double* ru,
                                                         double* chat, double* lhat, double* uhat,
                                                         double* rlhat, double* ruhat) {
  const int64_t most = n > m ? n : m;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < most;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    if (i < n) {
      const double d = dc[i], inv = 1.0 / d;
      chat[i] = cost[i] * d;
      lhat[i] = finite_bound(cl[i]) ? cl[i] * inv : cl[i];
      uhat[i] = finite_bound(cu[i]) ? cu[i] * inv : cu[i];
    }
    if (i < m) {
      const double d = dr[i];
      rlhat[i] = finite_bound(rl[i]) ? rl[i] * d : rl[i];
      ruhat[i] = finite_bound(ru[i]) ? ru[i] * d : ru[i];
    }
  }
}

}  // namespace

}  // namespace sankhya_hip
