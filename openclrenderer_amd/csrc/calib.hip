// calib.hip — a plain device-to-device copy of known size.  Profiling runs
// read its FETCH_SIZE / WRITE_SIZE counters to scale the counters of the
// path's kernels into HBM bytes.  Not part of the rendering ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ void __launch_bounds__(256) k_calib_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" int ocr_calib_copy(size_t bytes, int iters) {
  size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return -1;
  if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); return -1; }
  if (hipMemset(a, 1, n * 16) != hipSuccess) return -2;
  for (int i = 0; i < iters; i++)
    hipLaunchKernelGGL(k_calib_copy, dim3(4096), dim3(256), 0, nullptr, (const uint4*)a, (uint4*)b, n);
  hipError_t e = hipDeviceSynchronize();
  (void)hipFree(a);
  (void)hipFree(b);
  return e == hipSuccess ? 0 : -3;
}
