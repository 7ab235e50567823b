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

// Device scalar slots used by the Krylov kernels.  Slots written together by
// one two-dot reduction are consecutive (kScTrue/kScTrueS, kScTrue2/kScTrue2S).
enum ScIdx : int {
  kScRho = 0,
  kScRhoPrev = 1,
  kScBeta = 2,
  kScAlpha = 3,
  kScGamma = 4,
  kScRR = 5,       // ||r||^2 of the (inner) recursion
  kScRR0 = 6,
  kScTrue = 7,     // ||f - K u||^2 (FP64 true residual) ...
  kScTrueS = 8,    // ... and ||S (f - K u)||^2
  kScF2 = 9,       // ||f||^2
  kScInnerN = 10,  // ||r_hat||^2 of the current inner right-hand side
  kScTol2 = 11,    // convergence threshold on kScRR
  kScTrue2 = 15,   // ||f - K u||^2 of the restart refinement loop ...
  kScTrue2S = 16,  // ... and its scaled companion
  kScTmp = 17,
  kScCount = 24,
};

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

// Deterministic grid reductions, in two launches.  A reducing kernel stores
// one partial per block (thread 0 of block b writes partials[k*nb + b]); a
// one-block finalise kernel (dot_finalize, blas.hip) then sums the partials
// in index order and applies the solver bookkeeping.  No device-scope fences
// in the reducing kernel: on a multi-XCD GPU each such fence writes back the
// non-coherent L2, which costs far more than one extra small launch.
template <int K>
__device__ __forceinline__ void store_block_partials(const double (&vals)[K], double* partials) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = vals[k];
  }
}

// Bookkeeping applied by the finalise kernel to the reduced sums.
enum FinMode : int {
  kFinStore = 0,  // out[k] = sum_k
  kFinRho = 1,    // rho: sc[Rho] = sum; sc[Beta] = first ? 0 : rho / rho_prev; sc[RhoPrev] = rho
  kFinGamma = 2,  // gamma: sc[Gamma] = sum; sc[Alpha] = rho / gamma
  kFinRR = 3,     // ||r||^2: sc[RR] = sum; count the iteration; done flag when sc[RR] <= sc[Tol2]
  kFinRhoFlex = 4,  // rho and z.q_old: flexible (Polak-Ribiere) beta = -alpha (z.q_old) / rho_prev
};

}  // namespace wn
