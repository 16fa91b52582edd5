// copy_probe: a plain copy kernel of known size, y = x over 2^27 doubles (1 GiB read, 1 GiB
// written), launched three times. Profiling reference: what rocprofv3's FETCH_SIZE and
// WRITE_SIZE report for it calibrates the same counters on the engine's kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_kernel(const double* __restrict__ x, double* __restrict__ y, long n) {
  long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) y[i] = x[i];
}

int main() {
  const long n = 1L << 27;
  double *x = nullptr, *y = nullptr;
  if (hipMalloc(&x, n * 8) != hipSuccess || hipMalloc(&y, n * 8) != hipSuccess) return 1;
  if (hipMemset(x, 0, n * 8) != hipSuccess) return 1;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
  for (int r = 0; r < 3; ++r) {
    (void)hipEventRecord(a);
    copy_kernel<<<8192, 256>>>(x, y, n);
    (void)hipEventRecord(b);
    if (hipEventSynchronize(b) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    std::printf("copy %d: %.3f ms, %.1f GB/s (read+write)\n", r, ms, 2.0 * n * 8 / (ms * 1e6));
  }
  (void)hipFree(x);
  (void)hipFree(y);
  return 0;
}
