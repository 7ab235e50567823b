// Device layout of block-sparse matrices: sliced ELL over block rows.
//
// Block rows are grouped into slices of 64 (one wavefront lane per block
// row).  Rows are first sorted by block count inside windows of `sigma`
// consecutive rows (stable, longest first), so the rows sharing a slice have
// nearly equal lengths; the slice is then padded to its longest row.  Within
// a slice, block j of the row in lane l sits at slot  off[s] + j*64 + l,  and
// entry k (row-major inside the BRxBC block) of that block sits at value
// index  off[s]*BB + (j*BB + k)*64 + l.  A wavefront therefore reads every
// block entry of 64 rows as one contiguous 256 B (FP32) / 512 B (FP64) load,
// and each lane accumulates whole block rows in registers: no cross-lane
// reduction and no atomics.  Padding lanes have length 0 and row -1.
#pragma once
#include "common.hpp"

namespace wn {

template <typename T>
struct SellView {
  int n_rows;              // block rows
  int n_slices;
  const long long* off;    // [n_slices+1] slot offset of each slice
  const int* lane_len;     // [n_slices*64]
  const int* lane_row;     // [n_slices*64]
  const int* col;          // [slots] block column
  const T* val;            // [slots*BB]
};

template <typename T> struct AccT { using type = T; };

// Slice handled by this block.  With xcd_remap the 8 XCDs (blocks are issued
// round-robin over XCDs) each get one contiguous range of slices, so rows that
// share columns of x also share an L2.
__device__ __forceinline__ int slice_of_block(int n_slices, int xcd_remap) {
  const int b = blockIdx.x;
  if (!xcd_remap) return b;
  const int per = (n_slices + 7) >> 3;
  const int xcd = b & 7, k = b >> 3;
  return xcd * per + k;  // may exceed n_slices - 1 for the last XCD: caller checks
}

// Generic sliced-ELL block SpMV.  W wavefronts share one slice and split its
// block columns round-robin; their partial sums meet in LDS.  `Op` receives
// (block row, accumulated BR values) for every real row and may return a
// contribution to one fused dot product, reduced deterministically.
template <int BR, int BC, typename VT, typename XT, int W, class Op>
__global__ __launch_bounds__(64 * W) void sell_spmv_kernel(SellView<VT> A, const XT* __restrict__ x, Op op,
                                                           int xcd_remap, const int* skip) {
  if (skip && *skip) return;  // uniform: every block of the launch sees the same flag
  using AT = typename AccT<VT>::type;
  constexpr int BB = BR * BC;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int s = slice_of_block(A.n_slices, xcd_remap);  // uniform over the block
  double dot = 0.0;
  if (s < A.n_slices) {
    const int slot_lane = s * 64 + lane;
    const int len = A.lane_len[slot_lane];
    const long long base = A.off[s];
    AT acc[BR];
#pragma unroll
    for (int r = 0; r < BR; ++r) acc[r] = AT(0);
    const int* __restrict__ colp = A.col + base + lane;
    const VT* __restrict__ valp = A.val + base * BB + lane;
    for (int j = wave; j < len; j += W) {
      const int c = colp[(long long)j * 64];
      const VT* v = valp + (long long)j * BB * 64;
      XT xv[BC];
#pragma unroll
      for (int q = 0; q < BC; ++q) xv[q] = x[(long long)c * BC + q];
#pragma unroll
      for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int q = 0; q < BC; ++q) acc[r] = fma(AT(v[(r * BC + q) * 64]), AT(xv[q]), acc[r]);
    }
    if constexpr (W > 1) {
      __shared__ AT red[W - 1][BR][64];
      if (wave > 0) {
#pragma unroll
        for (int r = 0; r < BR; ++r) red[wave - 1][r][lane] = acc[r];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int w = 0; w < W - 1; ++w)
#pragma unroll
          for (int r = 0; r < BR; ++r) acc[r] += red[w][r][lane];
      }
    }
    if (wave == 0 && len > 0) dot = op.row(A.lane_row[slot_lane], acc);
  }
  if constexpr (Op::kDot) {
    __shared__ double sred[W];
    const double bs = block_sum(dot, sred);
    double v[1] = {bs};
    if (grid_reduce_last<1>(v, op.partials, op.counter, op.dot_out)) op.finalize();
  }
}

// ---------------------------------------------------------------- epilogues
// y = alpha * A x + beta * y
template <int BR, typename YT, typename AT>
struct OpAxpby {
  static constexpr bool kDot = false;
  YT* y;
  AT alpha, beta;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* dot_out = nullptr;
  __device__ __forceinline__ double row(int i, const AT (&acc)[BR]) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      const long long k = (long long)i * BR + r;
      AT v = alpha * acc[r];
      if (beta != AT(0)) v += beta * AT(y[k]);
      y[k] = YT(v);
    }
    return 0.0;
  }
  __device__ void finalize() const {}
};

// y = A p and gamma = p . (A p); finalize: alpha = rho / gamma (PCG step length).
template <int BR, typename YT, typename AT>
struct OpSpmvDot {
  static constexpr bool kDot = true;
  YT* y;
  const YT* p;        // same vector as the gathered x (square matrix)
  double* partials;
  unsigned int* counter;
  double* dot_out;    // gamma
  const double* rho;  // may be null (plain fused dot)
  double* alpha;      // may be null
  __device__ __forceinline__ double row(int i, const AT (&acc)[BR]) const {
    double d = 0.0;
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      const long long k = (long long)i * BR + r;
      y[k] = YT(acc[r]);
      d += double(p[k]) * double(acc[r]);
    }
    return d;
  }
  __device__ void finalize() const {
    if (threadIdx.x == 0 && alpha) {
      const double g = *dot_out;
      *alpha = (g != 0.0) ? (*rho / g) : 0.0;
    }
  }
};

// r = b - A x (FP64 true residual) with ||r||^2, and optionally the scaled
// single-precision copy rs = scale_vec .* r used as the inner right-hand side.
template <int BR, typename AT>
struct OpResidual {
  static constexpr bool kDot = true;
  const double* b;
  double* r;           // may be null
  float* rs;           // may be null
  const double* scale; // per-DOF scale for rs (may be null = 1)
  double* partials;
  unsigned int* counter;
  double* dot_out;     // ||r||^2
  __device__ __forceinline__ double row(int i, const AT (&acc)[BR]) const {
    double d = 0.0;
#pragma unroll
    for (int r0 = 0; r0 < BR; ++r0) {
      const long long k = (long long)i * BR + r0;
      const double v = b[k] - double(acc[r0]);
      if (r) r[k] = v;
      if (rs) rs[k] = float(scale ? scale[k] * v : v);
      d += v * v;
    }
    return d;
  }
  __device__ void finalize() const {}
};

// One Chebyshev smoothing step on D^-1 A (D^-1 given as `dinv`):
//   res = b - A x_in;  d = c1 * d + c2 * dinv .* res;  x_out = x_in + d
template <int BR, typename T>
struct OpCheb {
  static constexpr bool kDot = false;
  const T* b;
  const T* dinv;
  const T* x_in;
  T* x_out;
  T* d;
  T c1, c2;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* dot_out = nullptr;
  __device__ __forceinline__ double row(int i, const T (&acc)[BR]) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      const long long k = (long long)i * BR + r;
      const T res = b[k] - acc[r];
      const T dn = c1 * d[k] + c2 * dinv[k] * res;
      d[k] = dn;
      x_out[k] = x_in[k] + dn;
    }
    return 0.0;
  }
  __device__ void finalize() const {}
};

// res = b - A x (working precision), written to `res`.
template <int BR, typename T>
struct OpDefect {
  static constexpr bool kDot = false;
  const T* b;
  T* res;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* dot_out = nullptr;
  __device__ __forceinline__ double row(int i, const T (&acc)[BR]) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      const long long k = (long long)i * BR + r;
      res[k] = b[k] - acc[r];
    }
    return 0.0;
  }
  __device__ void finalize() const {}
};

// x += A e (prolongation: the gathered vector is the coarse correction).
template <int BR, typename T>
struct OpAddTo {
  static constexpr bool kDot = false;
  T* x;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* dot_out = nullptr;
  __device__ __forceinline__ double row(int i, const T (&acc)[BR]) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) x[(long long)i * BR + r] += acc[r];
    return 0.0;
  }
  __device__ void finalize() const {}
};

}  // namespace wn
