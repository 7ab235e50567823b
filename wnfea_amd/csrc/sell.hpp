// Device layout of block-sparse matrices: sliced ELL over block rows.
//
// Each block row is owned by L lanes of a wavefront (L from the shape: 4 for
// the fine 3x3 operator, up to 64 for short matrices with long rows such as
// restriction and coarse operators), so a slice holds R = 64/L rows.  Rows
// are first sorted by block count inside windows of `sigma` consecutive rows
// (stable, longest first), so the rows sharing a slice have nearly equal
// lengths.  Block j of the row in slice position r goes to lane
// l = (j mod L)*R + r at step t = j/L; its slot is off[s] + t*64 + l.  The
// sub-lane index is the high part of the lane number, so in the last, partial
// step of a slice the active lanes form one contiguous run and the fetched
// cache lines carry no idle lanes' padding.  The BB entries of the 64 blocks of one step form a
// region of 64*BB values starting at (off[s] + t*64)*BB, cut into chunks of
// V = 16 bytes / sizeof(value) entries: chunk g of lane l holds entries
// g*V .. g*V+V-1 (row-major inside the block) at region + g*64*V + l*V, and
// the BB mod V entries left over sit at region + (BB/V)*64*V + l*(BB mod V).
// A lane thus reads a block with 16-byte loads (three for an FP32 3x3
// block), and each load instruction of a wavefront covers 1 KiB of
// contiguous memory.  Each lane accumulates whole block rows in registers and
// the L lanes of a row meet in a shuffle reduction.  No atomics.  A lane
// derives its step count from its row's block count (one 8-byte record per
// slice row, shared by the row's L lanes); padding rows have row -1 and no
// blocks.
#pragma once
#include "common.hpp"

namespace wn {

template <typename T>
struct SellView {
  int n_rows;              // block rows
  int n_slices;
  int lanes;               // L: lanes per block row (power of two, <= 64)
  const long long* off;    // [n_slices+1] slot offset of each slice
  const int2* rinfo;       // [n_slices*64/lanes] {block row (-1: none), blocks in the row}
  const int* col;          // [slots] block column
  const T* val;            // [slots*BB]
};

// Steps of sub-lane q of a row with `blocks` blocks over L lanes.
__device__ __forceinline__ int lane_steps(int blocks, int q, int L) {
  return blocks > q ? (blocks - q + L - 1) / L : 0;
}

// Index of entry e (row-major) of the block in `slot` in a VT value array.
template <typename VT>
__host__ __device__ __forceinline__ long long val_index(long long slot, int e, int bb) {
  constexpr int V = 16 / (int)sizeof(VT);
  const long long region = (slot & ~63LL) * bb;
  const int l = (int)(slot & 63);
  const int g = e / V, full = bb / V;
  if (g < full) return region + (long long)g * 64 * V + l * V + (e - g * V);
  return region + (long long)full * 64 * V + l * (bb - full * V) + (e - full * V);
}

template <typename VT, int N>
struct alignas(N * sizeof(VT)) VecN {
  VT e[N];
};

#ifndef WN_NT_VALUES
#define WN_NT_VALUES 1
#endif

// N consecutive values (one load instruction).  With WN_NT_VALUES (the
// default) the load is marked non-temporal: matrix values and column indices
// are read once per product, so they should not displace x from L2 and the
// Infinity Cache (measured on the I-beam matrix: -9 % SpMV time in FP32 and
// FP64).
template <typename VT, int N>
__device__ __forceinline__ void load_vec(const VT* __restrict__ p, VT* out) {
#if WN_NT_VALUES
  if constexpr (N == 1) {
    out[0] = __builtin_nontemporal_load(p);
  } else {
    using E = VT __attribute__((ext_vector_type(N)));
    const E e = __builtin_nontemporal_load(reinterpret_cast<const E*>(p));
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = e[i];
  }
#else
  const VecN<VT, N> c = *reinterpret_cast<const VecN<VT, N>*>(p);
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = c.e[i];
#endif
}

__device__ __forceinline__ int load_col(const int* __restrict__ p) {
#if WN_NT_VALUES
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

// The BB entries of this lane's block in the region of one step.
template <typename VT, int BB>
__device__ __forceinline__ void load_block(const VT* __restrict__ region, int lane, VT (&v)[BB]) {
  constexpr int V = 16 / (int)sizeof(VT);
  constexpr int FULL = BB / V, REM = BB - FULL * V;
  static_assert((REM & (REM - 1)) == 0, "leftover entries must form a power-of-two load");
#pragma unroll
  for (int g = 0; g < FULL; ++g) load_vec<VT, V>(region + g * 64 * V + lane * V, v + g * V);
  if constexpr (REM > 0) load_vec<VT, REM>(region + FULL * 64 * V + lane * REM, v + FULL * V);
}

template <typename T> struct AccT { using type = T; };
template <> struct AccT<_Float16> { using type = float; };  // FP16 storage, FP32 arithmetic

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
// Launch guard read from device flags (uniform over the launch):
//   mode 0: always run; 1: skip once flags[0] (done) is set.
__device__ __forceinline__ bool guard_skip(const int* flags, int mode) {
  return flags && mode != 0 && flags[0] != 0;
}

// Generic sliced-ELL block SpMV.  W wavefronts share one slice and split its
// steps round-robin; their partial sums meet in LDS.  Each lane keeps U blocks
// in flight per loop trip (U column loads, then U*BB value loads and U x
// gathers, then the FMAs) to cover memory latency.  `Op` receives (block row,
// accumulated BR values) for every real row and may add to up to two fused
// dot products, reduced deterministically.
template <int BR, int BC, typename VT, typename XT, int W, int U, class Op>
__global__ __launch_bounds__(64 * W) void sell_spmv_kernel(SellView<VT> A, const XT* __restrict__ x, Op op,
                                                           int xcd_remap, const int* flags, int guard) {
  if (guard_skip(flags, guard)) return;
  using AT = typename AccT<VT>::type;
  constexpr int BB = BR * BC;
  // Loads per block and lane: the column index, the value chunks and the x
  // gathers (at most one per entry).  For blocks of more than 96 bytes the
  // loop bodies below ask the scheduler to issue all of a trip's loads
  // before its FMAs, so that they are in flight together: left to itself it
  // interleaves them with the FMAs to save registers and waits on each one
  // in turn (six or more serial waits per FP32 6x6 block).  Smaller blocks
  // are scheduled well without the hint.
  constexpr int kLoadsPerBlock = 1 + (BB * (int)sizeof(VT) + 15) / 16 + BC;
  constexpr bool kWideBlock = BB * (int)sizeof(VT) > 96;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int s = slice_of_block(A.n_slices, xcd_remap);  // uniform over the block
  double dot[2] = {0.0, 0.0};
  if (s < A.n_slices) {
    const int rps = 64 / A.lanes;  // rows per slice; lane = q * rps + r
    const int2 ri = A.rinfo[(long long)s * rps + lane % rps];
    const int q = lane / rps;
    const int len = lane_steps(ri.y, q, A.lanes);  // steps of this lane
    const long long base = A.off[s];
    // With at least BR lanes per row each output has its own lane (output q
    // on sub-lane q): its epilogue operands are loaded now, so that their
    // latency hides behind the block stream instead of trailing it.
    const bool one_out = A.lanes >= BR;
    const bool has_out = wave == 0 && ri.x >= 0 && q < BR;
    typename Op::Pre pre{};
    if (one_out && has_out) op.load((long long)ri.x * BR + q, pre);
    AT acc[BR];
#pragma unroll
    for (int r = 0; r < BR; ++r) acc[r] = AT(0);
    const int* __restrict__ colp = A.col + base + lane;
    const VT* __restrict__ valp = A.val + base * BB;  // region of step 0
    int j = wave;
    if constexpr (U > 1) {
      for (; j + (U - 1) * W < len; j += U * W) {
        int c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = load_col(colp + (long long)(j + u * W) * 64);
        VT v[U][BB];
#pragma unroll
        for (int u = 0; u < U; ++u) load_block<VT, BB>(valp + (long long)(j + u * W) * 64 * BB, lane, v[u]);
        XT xv[U][BC];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int q = 0; q < BC; ++q) xv[u][q] = x[(long long)c[u] * BC + q];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int q = 0; q < BC; ++q) acc[r] = fma(AT(v[u][r * BC + q]), AT(xv[u][q]), acc[r]);
        if constexpr (kWideBlock) {
          __builtin_amdgcn_sched_group_barrier(0x020, kLoadsPerBlock * U, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, 4 * BB * U, 0);
        }
      }
    }
    for (; j < len; j += W) {
      const int c = load_col(colp + (long long)j * 64);
      VT v[BB];
      load_block<VT, BB>(valp + (long long)j * 64 * BB, lane, v);
      XT xv[BC];
#pragma unroll
      for (int q = 0; q < BC; ++q) xv[q] = x[(long long)c * BC + q];
#pragma unroll
      for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int q = 0; q < BC; ++q) acc[r] = fma(AT(v[r * BC + q]), AT(xv[q]), acc[r]);
      if constexpr (kWideBlock) {
        __builtin_amdgcn_sched_group_barrier(0x020, kLoadsPerBlock, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 4 * BB, 0);
      }
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
    // the L lanes of a row (lane numbers r, r + R, r + 2R, ...) hold partial
    // sums: butterfly over the high bits of the lane number
    for (int off = 32; off >= rps; off >>= 1)
#pragma unroll
      for (int r = 0; r < BR; ++r) acc[r] += __shfl_xor(acc[r], off, 64);
    const int row = ri.x;
    // every lane of the row holds the full sums after the butterfly: the row's
    // L lanes share its BR outputs (output r on sub-lane r mod L)
    if (one_out) {
      if (has_out) {
        AT a = acc[0];
#pragma unroll
        for (int r = 1; r < BR; ++r)
          if (q == r) a = acc[r];
        op.out((long long)row * BR + q, a, dot, pre);
      }
    } else if (wave == 0 && row >= 0) {
      op.row(row, acc, dot, q, A.lanes);
    }
  }
  if constexpr (Op::kDots > 0) {
    __shared__ double sred[W];
    double v[Op::kDots];
#pragma unroll
    for (int k = 0; k < Op::kDots; ++k) v[k] = block_sum(dot[k], sred);
    store_block_partials<Op::kDots>(v, op.partials);
  }
}

// ---------------------------------------------------------------- epilogues
// y = alpha * A x + beta * y
template <int BR, typename YT, typename AT>
struct OpAxpby {
  static constexpr int kDots = 0;
  YT* y;
  AT alpha, beta;
  double* partials = nullptr;
  struct Pre { AT y; };
  __device__ __forceinline__ void load(long long k, Pre& p) const { p.y = beta != AT(0) ? AT(y[k]) : AT(0); }
  __device__ __forceinline__ void out(long long k, AT a, double (&)[2], const Pre& p) const {
    y[k] = YT(alpha * a + beta * p.y);
  }
  __device__ __forceinline__ void row(int i, const AT (&acc)[BR], double (&)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      if ((r & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r;
      AT v = alpha * acc[r];
      if (beta != AT(0)) v += beta * AT(y[k]);
      y[k] = YT(v);
    }
  }
};

// y = R x (restriction) and, from it, the coarse level's first Chebyshev
// iterate from x = 0: x1 = d1 = c2 dinv .* y (the expression of the
// separate first-step kernel, so the bits are the same).
template <int BR, typename T, typename AT>
struct OpRestrictFirst {
  static constexpr int kDots = 0;
  T* y;
  const T* dinv;
  T* x1;
  T* d1;
  T c2;
  double* partials = nullptr;
  struct Pre { T dinv; };
  __device__ __forceinline__ void load(long long k, Pre& p) const { p.dinv = dinv[k]; }
  __device__ __forceinline__ void put(long long k, T b, T dv) const {
    y[k] = b;
    const T v = c2 * dv * b;
    d1[k] = v;
    x1[k] = v;
  }
  __device__ __forceinline__ void out(long long k, AT a, double (&)[2], const Pre& p) const { put(k, T(a), p.dinv); }
  __device__ __forceinline__ void row(int i, const AT (&acc)[BR], double (&)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      if ((r & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r;
      put(k, T(acc[r]), dinv[k]);
    }
  }
};

// y = A p and gamma = p . (A p) (block partials; the finaliser forms alpha = rho / gamma).
template <int BR, typename YT, typename AT>
struct OpSpmvDot {
  static constexpr int kDots = 1;
  YT* y;
  const YT* p;        // same vector as the gathered x (square matrix)
  double* partials;
  struct Pre { double p; };
  __device__ __forceinline__ void load(long long k, Pre& pr) const { pr.p = double(p[k]); }
  __device__ __forceinline__ void out(long long k, AT a, double (&d)[2], const Pre& pr) const {
    y[k] = YT(a);
    d[0] += pr.p * double(a);
  }
  __device__ __forceinline__ void row(int i, const AT (&acc)[BR], double (&d)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      if ((r & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r;
      y[k] = YT(acc[r]);
      d[0] += double(p[k]) * double(acc[r]);
    }
  }
};

// r = b - A x (FP64 true residual): dots ||r||^2 and ||scale .* r||^2; writes
// r (optional) and the single-precision rs = scale .* r (optional).
template <int BR, typename AT>
struct OpResidual {
  static constexpr int kDots = 2;
  const double* b;
  double* r;            // may be null
  float* rs;            // may be null
  const double* scale;  // per-DOF scale (may be null = 1)
  double* partials;     // [0] ||r||^2, [1] ||scale r||^2 partials
  struct Pre { double b, s; };
  __device__ __forceinline__ void load(long long k, Pre& p) const {
    p.b = b[k];
    p.s = scale ? scale[k] : 1.0;
  }
  __device__ __forceinline__ void out(long long k, AT a, double (&d)[2], const Pre& p) const {
    const double v = p.b - double(a);
    const double vs = p.s * v;
    if (r) r[k] = v;
    if (rs) rs[k] = float(vs);
    d[0] += v * v;
    d[1] += vs * vs;
  }
  __device__ __forceinline__ void row(int i, const AT (&acc)[BR], double (&d)[2], int q, int L) const {
#pragma unroll
    for (int r0 = 0; r0 < BR; ++r0) {
      if ((r0 & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r0;
      const double v = b[k] - double(acc[r0]);
      const double vs = scale ? scale[k] * v : v;
      if (r) r[k] = v;
      if (rs) rs[k] = float(vs);
      d[0] += v * v;
      d[1] += vs * vs;
    }
  }
};

// One Chebyshev smoothing step on D^-1 A (D^-1 given as `dinv`):
//   res = b - A x_in;  d = c1 * d + c2 * dinv .* res;  x_out = x_in + d
// Unit: D = I (an equilibrated operator's unit diagonal); dinv is not read.
template <int BR, typename T, bool Unit = false>
struct OpCheb {
  static constexpr int kDots = 0;
  const T* b;
  const T* dinv;
  const T* x_in;
  T* x_out;
  T* d;
  T c1, c2;
  double* partials = nullptr;
  struct Pre { T b, dinv, x, d; };
  __device__ __forceinline__ void load(long long k, Pre& p) const {
    p.b = b[k];
    if constexpr (Unit) {
      p.dinv = T(1);
    } else {
      p.dinv = dinv[k];
    }
    p.x = x_in[k];
    p.d = d[k];
  }
  __device__ __forceinline__ void out(long long k, T a, double (&)[2], const Pre& p) const {
    const T dn = c1 * p.d + c2 * p.dinv * (p.b - a);
    d[k] = dn;
    x_out[k] = p.x + dn;
  }
  __device__ __forceinline__ void row(int i, const T (&acc)[BR], double (&)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      if ((r & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r;
      const T res = b[k] - acc[r];
      const T dn = c1 * d[k] + c2 * (Unit ? T(1) : dinv[k]) * res;
      d[k] = dn;
      x_out[k] = x_in[k] + dn;
    }
  }
};

// res = b - A x (working precision), written to `res`.
template <int BR, typename T>
struct OpDefect {
  static constexpr int kDots = 0;
  const T* b;
  T* res;
  double* partials = nullptr;
  struct Pre { T b; };
  __device__ __forceinline__ void load(long long k, Pre& p) const { p.b = b[k]; }
  __device__ __forceinline__ void out(long long k, T a, double (&)[2], const Pre& p) const { res[k] = p.b - a; }
  __device__ __forceinline__ void row(int i, const T (&acc)[BR], double (&)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      if ((r & (L - 1)) != q) continue;
      const long long k = (long long)i * BR + r;
      res[k] = b[k] - acc[r];
    }
  }
};

// x += A e (prolongation: the gathered vector is the coarse correction).
template <int BR, typename T>
struct OpAddTo {
  static constexpr int kDots = 0;
  T* x;
  double* partials = nullptr;
  struct Pre { T x; };
  __device__ __forceinline__ void load(long long k, Pre& p) const { p.x = x[k]; }
  __device__ __forceinline__ void out(long long k, T a, double (&)[2], const Pre& p) const { x[k] = p.x + a; }
  __device__ __forceinline__ void row(int i, const T (&acc)[BR], double (&)[2], int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r)
      if ((r & (L - 1)) == q) x[(long long)i * BR + r] += acc[r];
  }
};

}  // namespace wn
