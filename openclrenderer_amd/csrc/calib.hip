// calib.hip — a plain device-to-device copy of known size.  Profiling runs
// read its FETCH_SIZE / WRITE_SIZE counters to scale the counters of the
// path's kernels into HBM bytes.  Not part of the rendering ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ocr_math.h"

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

// Atomic throughput probes: n u32 atomicMin (or u64 atomicMax) on addresses
// drawn from a hash of the work-item index over `words` targets; `cluster`
// consecutive work items share a 64-address window (1 = uniform random).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(256) k_calib_atomic_min(uint32_t* __restrict__ buf, uint32_t words, size_t n,
                                                          uint32_t cluster) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t g = (uint32_t)(i / cluster);
    uint32_t a = cluster == 1 ? mix32((uint32_t)i) % words
                              : (mix32(g) % (words / 64)) * 64 + (uint32_t)(i % 64);
    atomicMin(&buf[a], mix32((uint32_t)i ^ 0x9e3779b9u));
  }
}

__global__ void __launch_bounds__(256) k_calib_atomic_max64(unsigned long long* __restrict__ buf, uint32_t words,
                                                            size_t n, uint32_t cluster) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t g = (uint32_t)(i / cluster);
    uint32_t a = cluster == 1 ? mix32((uint32_t)i) % words
                              : (mix32(g) % (words / 64)) * 64 + (uint32_t)(i % 64);
    atomicMax(&buf[a], (unsigned long long)mix32((uint32_t)i));
  }
}

__global__ void __launch_bounds__(256) k_calib_scatter_store(uint32_t* __restrict__ buf, uint32_t words, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) buf[mix32((uint32_t)i) % words] = (uint32_t)i;
}

// As k_calib_atomic_min at workgroup scope: the atomic is performed in the
// issuing XCD's L2 (throughput probe only; its values are meaningless when
// several XCDs touch the buffer).
__global__ void __launch_bounds__(256) k_calib_atomic_min_wg(uint32_t* __restrict__ buf, uint32_t words, size_t n,
                                                             uint32_t cluster) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint32_t g = (uint32_t)(i / cluster);
    uint32_t a = cluster == 1 ? mix32((uint32_t)i) % words
                              : (mix32(g) % (words / 64)) * 64 + (uint32_t)(i % 64);
    __hip_atomic_fetch_min(&buf[a], mix32((uint32_t)i ^ 0x9e3779b9u), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// One 64-bit atomicAdd per wavefront on a single address (n wavefront
// iterations), the pattern of a global work-list counter.
__global__ void __launch_bounds__(256) k_calib_atomic_same(unsigned long long* __restrict__ counter, size_t n) {
  const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const size_t nwaves = (size_t)gridDim.x * blockDim.x / 64;
  for (size_t i = wave; i < n; i += nwaves)
    if ((threadIdx.x & 63) == 0) atomicAdd(counter, 1ull);
}

// Returns the average kernel time in microseconds of `iters` launches
// (mode 0: u32 atomicMin, 1: u64 atomicMax, 2: plain scattered store,
// 3: one same-address u64 atomicAdd per wavefront, n wavefront iterations,
// 4: u32 atomicMin at workgroup scope).
extern "C" double ocr_calib_atomics(int mode, uint32_t words, size_t n, uint32_t cluster, int iters) {
  void* buf = nullptr;
  if (hipMalloc(&buf, (size_t)words * 8) != hipSuccess) return -1.0;
  if (hipMemset(buf, 0xFF, (size_t)words * 8) != hipSuccess) return -2.0;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  dim3 grid(4096), block(256);
  for (int w = 0; w < 2 + iters; w++) {
    if (w == 2) (void)hipEventRecord(e0, nullptr);
    if (mode == 0) hipLaunchKernelGGL(k_calib_atomic_min, grid, block, 0, nullptr, (uint32_t*)buf, words, n, cluster);
    else if (mode == 1) hipLaunchKernelGGL(k_calib_atomic_max64, grid, block, 0, nullptr, (unsigned long long*)buf, words, n, cluster);
    else if (mode == 2) hipLaunchKernelGGL(k_calib_scatter_store, grid, block, 0, nullptr, (uint32_t*)buf, words, n);
    else if (mode == 4) hipLaunchKernelGGL(k_calib_atomic_min_wg, grid, block, 0, nullptr, (uint32_t*)buf, words, n, cluster);
    else hipLaunchKernelGGL(k_calib_atomic_same, grid, block, 0, nullptr, (unsigned long long*)buf, n);
  }
  (void)hipEventRecord(e1, nullptr);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  return e == hipSuccess ? (double)ms * 1e3 / iters : -3.0;
}

// VALU issue ceiling: every work item runs `iters` rounds of 8 independent
// FMA chains (mode 0: binary32 v_fma_f32, mode 1: binary64 v_fma_f64), so
// each round issues 8 VALU instructions per wavefront with no dependency
// stall; the grid holds 8 wavefronts per SIMD.  The sums are stored so the
// chains stay live.
template <typename T>
__global__ void __launch_bounds__(256) k_calib_valu(T* __restrict__ out, uint32_t iters, T a, T b) {
#pragma clang fp contract(off)
  T x[8];
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = (T)(threadIdx.x + k);
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if constexpr (sizeof(T) == 4) x[k] = __builtin_fmaf(x[k], a, b);
      else x[k] = __builtin_fma(x[k], a, b);
    }
  }
  T s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s += x[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Returns the measured VALU rate in wave-instructions per second (the FMAs
// only; loop control is scalar), or a negative value on failure.
extern "C" double ocr_calib_valu(int mode, uint32_t iters) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1.0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1.0;
  const uint32_t blocks = (uint32_t)cus * 4u * 8u / 4u;   // 8 wavefronts per SIMD, 4 per block
  void* buf = nullptr;
  if (hipMalloc(&buf, (size_t)blocks * 256 * 8) != hipSuccess) return -1.0;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int w = 0; w < 3; w++) {
    if (w == 2) (void)hipEventRecord(e0, nullptr);
    if (mode == 0)
      hipLaunchKernelGGL(k_calib_valu<float>, dim3(blocks), dim3(256), 0, nullptr, (float*)buf, iters, 0.999f, 0.001f);
    else
      hipLaunchKernelGGL(k_calib_valu<double>, dim3(blocks), dim3(256), 0, nullptr, (double*)buf, iters, 0.999, 0.001);
  }
  (void)hipEventRecord(e1, nullptr);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  if (e != hipSuccess || ms <= 0.f) return -3.0;
  const double wave_insts = (double)blocks * 4.0 * (double)iters * 8.0;
  return wave_insts / ((double)ms * 1e-3);
}

// Evaluates one of the device's arithmetic-contract functions over n inputs
// (fn 0: exp_f), for comparison with the oracle's restatement.
__global__ void __launch_bounds__(256) k_calib_eval(int fn, const float* __restrict__ x, float* __restrict__ y,
                                                    uint32_t n) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = fn == 0 ? ocr::exp_f(x[i]) : x[i];
}

extern "C" int ocr_calib_eval(int fn, const float* x, float* y, uint32_t n) {
  float *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, (size_t)n * 4) != hipSuccess) return -1;
  if (hipMalloc(&dy, (size_t)n * 4) != hipSuccess) { (void)hipFree(dx); return -1; }
  hipError_t e = hipMemcpy(dx, x, (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_calib_eval, dim3((n + 255) / 256), dim3(256), 0, nullptr, fn, dx, dy, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, dy, (size_t)n * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return e == hipSuccess ? 0 : -3;
}

// Exhaustive checks of the short reciprocal sequences against IEEE
// arithmetic, over every binary32 bit pattern in [lo, hi] (NaN inputs must
// give NaN): mode 0 ocr::rcp_rn(b) vs 1.0f / b, mode 1 ocr::inv_sqrt_cr(d) vs
// 1.0f / sqrtf(d), mode 2 ocr::sqrt_cr(x) vs sqrtf(x).  One count per work item, added once, so a bad helper
// cannot turn the check into billions of same-address atomics.
__global__ void __launch_bounds__(256) k_calib_rcp_check(int mode, uint32_t lo, uint32_t hi,
                                                         unsigned long long* __restrict__ bad,
                                                         uint32_t* __restrict__ first_bad) {
#pragma clang fp contract(off)
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long n = 0;
  uint32_t first = 0xFFFFFFFFu;
  for (uint64_t u = (uint64_t)lo + i; u <= (uint64_t)hi; u += stride) {
    const float b = __uint_as_float((uint32_t)u);
    const float fast = mode == 0 ? ocr::rcp_rn(b) : mode == 1 ? ocr::inv_sqrt_cr(b) : ocr::sqrt_cr(b);
    const float ref = mode == 0 ? 1.0f / b : mode == 1 ? 1.0f / sqrtf(b) : sqrtf(b);
    const bool same = __float_as_uint(fast) == __float_as_uint(ref) || (ref != ref && fast != fast);
    if (!same) {
      n++;
      first = min(first, (uint32_t)u);
    }
  }
  if (n) {
    atomicAdd(bad, n);
    atomicMin(first_bad, first);
  }
}

// Candidate square-root sequences, compared bit for bit with the correctly
// rounded sqrtf over every binary32 pattern in [lo, hi]; mismatches are
// counted per input exponent (256 buckets).
//   mode 0: the hardware estimate, then the residual test on its two
//           neighbours (no scaling, no special-value handling);
//   mode 1: the hardware estimate corrected once by the residual times half
//           the hardware reciprocal square root.
__device__ __forceinline__ float sqrt_two_sided(float x) {
  float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __int_as_float(__float_as_int(s) - 1);
  const float su = __int_as_float(__float_as_int(s) + 1);
  const float rd = fmaf(-sd, s, x);
  const float ru = fmaf(-su, s, x);
  s = rd <= 0.f ? sd : s;
  s = ru > 0.f ? su : s;
  return s;
}
__device__ __forceinline__ float sqrt_rsq_step(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float h = 0.5f * __builtin_amdgcn_rsqf(x);
  const float r = fmaf(-s, s, x);
  return fmaf(r, h, s);
}

__global__ void __launch_bounds__(256) k_calib_sqrt_check(int mode, uint32_t lo, uint32_t hi,
                                                          unsigned long long* __restrict__ bad_by_exp) {
#pragma clang fp contract(off)
  __shared__ unsigned int local[256];
  for (int k = threadIdx.x; k < 256; k += blockDim.x) local[k] = 0;
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t u = (uint64_t)lo + i; u <= (uint64_t)hi; u += stride) {
    const float x = __uint_as_float((uint32_t)u);
    const float fast = mode == 0 ? sqrt_two_sided(x) : sqrt_rsq_step(x);
    const float ref = sqrtf(x);
    const bool same = __float_as_uint(fast) == __float_as_uint(ref) || (ref != ref && fast != fast);
    if (!same) atomicAdd(&local[((uint32_t)u >> 23) & 255], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 256; k += blockDim.x)
    if (local[k]) atomicAdd(&bad_by_exp[k], (unsigned long long)local[k]);
}

extern "C" int ocr_calib_sqrt_check(int mode, uint32_t lo, uint32_t hi, unsigned long long* bad_by_exp) {
  unsigned long long* d_bad = nullptr;
  if (hipMalloc(&d_bad, 256 * sizeof(*d_bad)) != hipSuccess) return -1;
  hipError_t e = hipMemset(d_bad, 0, 256 * sizeof(*d_bad));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_calib_sqrt_check, dim3(8192), dim3(256), 0, nullptr, mode, lo, hi, d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(bad_by_exp, d_bad, 256 * sizeof(*d_bad), hipMemcpyDeviceToHost);
  (void)hipFree(d_bad);
  return e == hipSuccess ? 0 : -3;
}

extern "C" int ocr_calib_rcp_check(int mode, uint32_t lo, uint32_t hi, unsigned long long* mismatches,
                                   uint32_t* first_bad) {
  unsigned long long* d_bad = nullptr;
  uint32_t* d_first = nullptr;
  if (hipMalloc(&d_bad, sizeof(*d_bad)) != hipSuccess) return -1;
  if (hipMalloc(&d_first, sizeof(*d_first)) != hipSuccess) { (void)hipFree(d_bad); return -1; }
  hipError_t e = hipMemset(d_bad, 0, sizeof(*d_bad));
  if (e == hipSuccess) e = hipMemset(d_first, 0xFF, sizeof(*d_first));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_calib_rcp_check, dim3(8192), dim3(256), 0, nullptr, mode, lo, hi, d_bad, d_first);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(mismatches, d_bad, sizeof(*d_bad), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(first_bad, d_first, sizeof(*d_first), hipMemcpyDeviceToHost);
  (void)hipFree(d_bad);
  (void)hipFree(d_first);
  return e == hipSuccess ? 0 : -3;
}
