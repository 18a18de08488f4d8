// backend.h -- the few runtime primitives the solver uses: device memory,
// copies, a 1-D parallel-for over a __host__ __device__ body, and FP64 dot
// products.
//
// The GPU build (hipcc, gfx950) maps them to HIP.  Defining FCBGA_HOST_EMU and
// compiling the same sources as C++ maps them to plain host loops: that build
// exists only so the test suite can check the kernels' indexing and the
// orchestration on a machine without a GPU (with the address sanitizer); the
// product library is always the HIP build.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>

#if defined(FCBGA_HOST_EMU)
#define FCBGA_GPU 0
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __global__
#define __global__
#endif
#ifndef __align__
#define __align__(n) alignas(n)
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#else
#define FCBGA_GPU 1
#include <hip/hip_runtime.h>
#endif

namespace be {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#if FCBGA_GPU
#define BE_CHECK(x)                                                                                   \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess)                                                                             \
      throw be::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + \
                      std::to_string(__LINE__));                                                      \
  } while (0)

typedef hipStream_t stream_t;
#else
typedef void *stream_t;
#endif

inline void *dmalloc(size_t bytes) {
  if (bytes == 0) bytes = 16;
#if FCBGA_GPU
  void *p = nullptr;
  BE_CHECK(hipMalloc(&p, bytes));
  return p;
#else
  void *p = calloc(1, bytes);
  if (!p) throw Error("host allocation failed");
  return p;
#endif
}

inline void dfree(void *p) {
  if (!p) return;
#if FCBGA_GPU
  (void)hipFree(p);
#else
  free(p);
#endif
}

inline void h2d(void *dst, const void *src, size_t bytes, stream_t s) {
  if (!bytes) return;
#if FCBGA_GPU
  BE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
#else
  (void)s;
  memcpy(dst, src, bytes);
#endif
}

inline void d2h(void *dst, const void *src, size_t bytes, stream_t s) {
  if (!bytes) return;
#if FCBGA_GPU
  BE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  BE_CHECK(hipStreamSynchronize(s));
#else
  (void)s;
  memcpy(dst, src, bytes);
#endif
}

inline void d2d(void *dst, const void *src, size_t bytes, stream_t s) {
  if (!bytes) return;
#if FCBGA_GPU
  BE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
#else
  (void)s;
  memmove(dst, src, bytes);
#endif
}

inline void sync(stream_t s) {
#if FCBGA_GPU
  BE_CHECK(hipStreamSynchronize(s));
#else
  (void)s;
#endif
}

#if FCBGA_GPU
template <typename F>
__global__ void __launch_bounds__(256) k_for(int64_t n, F f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f(i);
}
#endif

// parallel-for over [0, n); the body is a __host__ __device__ lambda
template <typename F>
inline void pfor(int64_t n, stream_t s, F f, int block = 256) {
  if (n <= 0) return;
#if FCBGA_GPU
  int64_t grid = (n + block - 1) / block;
  hipLaunchKernelGGL(k_for<F>, dim3((unsigned)grid), dim3(block), 0, s, n, f);
  BE_CHECK(hipGetLastError());
#else
  (void)s; (void)block;
  for (int64_t i = 0; i < n; i++) f(i);
#endif
}

// ---- FP64 dot products --------------------------------------------------------
// Two-pass reduction in a fixed order (deterministic for a given n): per
// block, each wavefront sums its lanes by a butterfly of lane exchanges and
// the block adds its four wavefront sums in order.

constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 1024;  // blocks of the first pass: partial sums per dot

#if FCBGA_GPU
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// the block's sum of v (RED_BLOCK threads), valid in thread 0; sm: RED_BLOCK / 64 doubles
__device__ __forceinline__ double block_sum(double v, double *sm) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < RED_BLOCK / 64; w++) t += sm[w];
  return t;
}

template <typename TA, typename TB>
__global__ void __launch_bounds__(RED_BLOCK) k_dot_partial(int64_t n, const TA *__restrict__ a,
                                                           const TB *__restrict__ b, double *__restrict__ part) {
  __shared__ double sm[RED_BLOCK / 64];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += (int64_t)RED_GRID * RED_BLOCK)
    s += (double)a[i] * (double)b[i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// one sum per block q: the n partials at part[q * n ...] into out[q * ostride]
__global__ void __launch_bounds__(RED_BLOCK) k_sum_partials(int n, const double *__restrict__ part,
                                                            double *__restrict__ out, int ostride) {
  __shared__ double sm[RED_BLOCK / 64];
  const int q = blockIdx.x;
  double s = 0;
  for (int i = threadIdx.x; i < n; i += RED_BLOCK) s += part[(int64_t)q * n + i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) out[(int64_t)q * ostride] = s;
}
#endif

// result[slot] = a . b  (device scalar); partials: RED_GRID doubles per slot of scratch
template <typename TA, typename TB>
inline void dot_async(int64_t n, const TA *a, const TB *b, double *partials, double *result, stream_t s) {
#if FCBGA_GPU
  hipLaunchKernelGGL((k_dot_partial<TA, TB>), dim3(RED_GRID), dim3(RED_BLOCK), 0, s, n, a, b, partials);
  BE_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_BLOCK), 0, s, RED_GRID, partials, result, 1);
  BE_CHECK(hipGetLastError());
#else
  (void)partials; (void)s;
  double acc = 0;
  for (int64_t i = 0; i < n; i++) acc += (double)a[i] * (double)b[i];
  *result = acc;
#endif
}

}  // namespace be
