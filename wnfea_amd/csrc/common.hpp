// Shared helpers of the gfx950 solve-path library: status codes, launch
// helpers, wavefront (64-lane) reductions and the deterministic "last block
// finalises" grid reduction used by every fused dot product.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace wn {

// Status convention of the C-ABI: 0 ok, <0 argument errors, >0 HIP errors.
enum Status : int {
  kOk = 0,
  kBadArg = -1,
  kBadSize = -2,
  kBadMode = -3,
  kNotReady = -4,
  kBadMesh = -5,
  kNoConverge = -6,
};

#define WN_HIP(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return static_cast<int>(_e); \
  } while (0)

constexpr int kWave = 64;

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum (blockDim.x a multiple of 64, at most 1024); result valid in
// every thread.  `smem` holds at least blockDim.x/64 values.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* smem) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < nw; ++w) t += smem[w];  // fixed order: deterministic
  return t;
}

// Deterministic grid reduction of K values.  Each block stores its partials
// at partials[k * gridDim.x + blockIdx.x]; the block that arrives last (one
// counter increment per block) sums every partial in index order and writes
// out[k].  The counter is left at zero for the next launch.  Returns true in
// the finalising block (after out[] is written and visible to that block).
template <int K>
__device__ bool grid_reduce_last(const double (&vals)[K], double* partials, unsigned int* counter,
                                 double* out) {
  __shared__ bool s_last;
  __shared__ double s_red[K][16];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = vals[k];
    __threadfence();
    const unsigned int t = atomicAdd(counter, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  const int nb = gridDim.x;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    // strided partial sums in fixed order, then fixed-order combine
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
      s += __hip_atomic_load(&partials[(size_t)k * nb + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s = wave_sum(s);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0) s_red[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += s_red[k][w];
      out[k] = t;
    }
    *counter = 0u;
    __threadfence();
  }
  __syncthreads();
  return true;
}

}  // namespace wn
