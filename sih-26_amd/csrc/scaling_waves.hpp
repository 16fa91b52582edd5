// scaling_waves.hpp - device pieces of the preconditioner (src/la/scaling.cpp:14-159) shared by
// the single-LP engine (pdhg_device.hip) and the batched engine (pdhg_batch.hip): one wave per
// matrix line for the Ruiz maxima, the scaling of the entries and the Pock-Chambolle sums in
// storage order, and the flat per-coordinate kernels (factors, fills, scaled vectors). The
// results equal the reference's host passes bit for bit: maxima do not depend on order, the
// sums are added in storage order as one chain, and factors and scaled values use the same
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

// Sum of |a| over [k0, k1) in storage order, one wave, as one sequential chain of adds, so the
// result is the sequential sum bit for bit. Each step's 64 kScU entries arrive with coalesced
// loads and are staged in `stage` (this wave's 64 kScU doubles of LDS); the loads of the next
// step are issued before the current step's entries are added. Every lane then reads the staged
// entries in storage order (one broadcast LDS read per entry) and runs the same chain, so only
// the dependent adds are on the critical path: on the densest row of the 200k x 500k Zipf LP
// (about 495,000 entries) reading the entries back with v_readlane instead took 6.9 ms.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ double wave_abssum_ordered(const double* val, int32_t k0, int32_t k1,
                                                      int lane, double* stage) {
  double r = 0.0;
  if (k1 <= k0) return r;
  // Loads are never conditional (a lane past the end reads entry k1 - 1 of this line and
  // multiplies it by 0.0, exact for the finite entries of a matrix), so all of a step's loads
  // stay in flight together instead of each waiting behind a branch.
  const int32_t last = k1 - 1;
  double v[kScU];
#pragma unroll
  for (int u = 0; u < kScU; ++u) {
    const int32_t k = k0 + u * 64 + lane;
    v[u] = fabs(val[k < last ? k : last]) * (k < k1 ? 1.0 : 0.0);
  }
  for (int32_t base = k0; base < k1; base += 64 * kScU) {
    const int32_t next = base + 64 * kScU;
    double nv[kScU];
#pragma unroll
    for (int u = 0; u < kScU; ++u) {
      const int32_t k = next + u * 64 + lane;
      nv[u] = fabs(val[k < last ? k : last]) * (k < k1 ? 1.0 : 0.0);
    }
#pragma unroll
    for (int u = 0; u < kScU; ++u) stage[u * 64 + lane] = v[u];  // entry base + i at stage[i]
    wave_sync_lds();
    const int32_t count = k1 - base < 64 * kScU ? k1 - base : 64 * kScU;
    // Groups of 16, the next group's reads issued before the current group's adds.
    const int32_t full = count & ~15;
    int32_t i = 0;
    if (full > 0) {
      double t[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) t[q] = stage[q];
#pragma unroll 1
      for (i = 16; i < full; i += 16) {
        double nt[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) nt[q] = stage[i + q];
#pragma unroll
        for (int q = 0; q < 16; ++q) r += t[q];
#pragma unroll
        for (int q = 0; q < 16; ++q) t[q] = nt[q];
      }
#pragma unroll
      for (int q = 0; q < 16; ++q) r += t[q];
    }
    for (; i < count; ++i) r += stage[i];
    wave_sync_lds();  // every lane has read the step before the next one is staged
#pragma unroll
    for (int u = 0; u < kScU; ++u) v[u] = nv[u];
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
