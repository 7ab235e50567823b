// Multiple right-hand sides: block SpMM, the V-cycle and PCG on NR columns
// at once.
//
// Vectors of NR columns are stored interleaved, X[dof][NR], so the NR values
// a block column gathers are contiguous (one or a few 16-byte loads) and each
// matrix block is read once for all NR columns: the matrix stream, which is
// ~90 % of an SpMV's bytes, is amortised NR ways.  Columns are independent
// CG runs (per-column rho, alpha, beta, ||r||^2 on the device); a column that
// has converged is frozen (alpha = beta = 0) while the others continue, and
// the loop ends when all are done.  Same precision split as the single-column
// solver: FP64 Krylov vectors, SpMM and updates; the V-cycle in FP32 on S K S
// as z = S M(S r).
#include <cmath>
#include "sell.hpp"
#include "solver.hpp"

namespace wn {

constexpr int kMrBlock = 256;

template <typename XT, int NR>
__device__ __forceinline__ void load_cols(const XT* __restrict__ p, XT (&v)[NR]) {
  constexpr int V = 16 / (int)sizeof(XT);
  static_assert(NR % V == 0, "columns must fill 16-byte loads");
#pragma unroll
  for (int g = 0; g < NR / V; ++g) {
    const VecN<XT, V> c = *reinterpret_cast<const VecN<XT, V>*>(p + g * V);
#pragma unroll
    for (int i = 0; i < V; ++i) v[g * V + i] = c.e[i];
  }
}

// Sliced-ELL block SpMM (layout of sell.hpp) on NR interleaved columns.
template <int BR, int BC, typename VT, typename XT, int W, int NR, class Op>
__global__ __launch_bounds__(64 * W) void sell_spmm_kernel(SellView<VT> A, const XT* __restrict__ x, Op op,
                                                           int xcd_remap, const int* flags, int guard) {
  if (guard_skip(flags, guard)) return;
  using AT = typename AccT<VT>::type;
  constexpr int BB = BR * BC;
  constexpr int kLoads = 1 + (BB * (int)sizeof(VT) + 15) / 16 + BC * ((NR * (int)sizeof(XT) + 15) / 16);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int s = slice_of_block(A.n_slices, xcd_remap);
  double dot[Op::kDots > 0 ? Op::kDots : 1];
#pragma unroll
  for (int k = 0; k < (Op::kDots > 0 ? Op::kDots : 1); ++k) dot[k] = 0.0;
  if (s < A.n_slices) {
    const int rps = 64 / A.lanes;  // rows per slice; a row's lanes are r, r + rps, ...
    const int2 ri = A.rinfo[(long long)s * rps + lane % rps];
    const int len = lane_steps(ri.y, lane / rps, A.lanes);
    const long long base = A.off[s];
    AT acc[BR * NR];
#pragma unroll
    for (int t = 0; t < BR * NR; ++t) acc[t] = AT(0);
    const int* __restrict__ colp = A.col + base + lane;
    const VT* __restrict__ valp = A.val + base * BB;
    for (int j = wave; j < len; j += W) {
      const int c = load_col(colp + (long long)j * 64);
      VT v[BB];
      load_block<VT, BB>(valp + (long long)j * 64 * BB, lane, v);
#pragma unroll
      for (int q = 0; q < BC; ++q) {
        XT xv[NR];
        load_cols<XT, NR>(x + ((long long)c * BC + q) * NR, xv);
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
          for (int k = 0; k < NR; ++k) acc[r * NR + k] = fma(AT(v[r * BC + q]), AT(xv[k]), acc[r * NR + k]);
      }
      // FP32 3x3 blocks: all of the block's loads (column index, values, the
      // NR-wide x rows) in flight together before its FMAs; left to itself
      // the compiler reuses the x registers and waits between the x loads
      // (I-beam, 8 columns: 277 -> 258 us).  The FP64 and 6x6 forms are left
      // to the compiler: the hint raised their registers and slowed the FP64
      // product (370 -> 467 us).
      if constexpr (sizeof(XT) == 4 && BB <= 9) {
        __builtin_amdgcn_sched_group_barrier(0x020, kLoads, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 2 * BB * NR, 0);
      }
    }
    if constexpr (W > 1) {
      __shared__ AT red[W - 1][BR * NR][64];
      if (wave > 0) {
#pragma unroll
        for (int t = 0; t < BR * NR; ++t) red[wave - 1][t][lane] = acc[t];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int w = 0; w < W - 1; ++w)
#pragma unroll
          for (int t = 0; t < BR * NR; ++t) acc[t] += red[w][t][lane];
      }
    }
    for (int off = 32; off >= rps; off >>= 1)
#pragma unroll
      for (int t = 0; t < BR * NR; ++t) acc[t] += __shfl_xor(acc[t], off, 64);
    const int row = ri.x;
    // every lane of the row holds the full sums after the butterfly: the row's
    // L lanes share the NR columns of the epilogue (column k on sub-lane k mod L)
    if (wave == 0 && row >= 0) op.row(row, acc, dot, lane / rps, A.lanes);
  }
  if constexpr (Op::kDots > 0) {
    __shared__ double sred[W];
    double v[Op::kDots];
#pragma unroll
    for (int k = 0; k < Op::kDots; ++k) v[k] = block_sum(dot[k], sred);
    store_block_partials<Op::kDots>(v, op.partials);
  }
}

// ---------------------------------------------------------------- epilogues (index (i*BR + r)*NR + k)
template <int BR, int NR, typename YT, typename AT>
struct MOpStore {  // y = A x, or y += A x with add
  static constexpr int kDots = 0;
  YT* y;
  int add;
  double* partials = nullptr;
  __device__ __forceinline__ void row(int i, const AT (&acc)[BR * NR], double*, int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if ((k & (L - 1)) != q) continue;
        const long long o = ((long long)i * BR + r) * NR + k;
        y[o] = add ? YT(AT(y[o]) + acc[r * NR + k]) : YT(acc[r * NR + k]);
      }
  }
};

template <int BR, int NR, typename T>
struct MOpSpmvDot {  // q = A p, gamma_k = p_k . q_k
  static constexpr int kDots = NR;
  T* y;
  const T* p;
  double* partials;
  __device__ __forceinline__ void row(int i, const T (&acc)[BR * NR], double* d, int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if ((k & (L - 1)) != q) continue;
        const long long o = ((long long)i * BR + r) * NR + k;
        y[o] = acc[r * NR + k];
        d[k] += double(p[o]) * double(acc[r * NR + k]);
      }
  }
};

template <int BR, int NR, typename T>
struct MOpCheb {  // res = b - A x_in; d = c1 d + c2 dinv res; x_out = x_in + d
  static constexpr int kDots = 0;
  const T* b;
  const T* dinv;  // per DOF
  const T* x_in;
  T* x_out;
  T* d;
  T c1, c2;
  double* partials = nullptr;
  __device__ __forceinline__ void row(int i, const T (&acc)[BR * NR], double*, int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      const T di = dinv[(long long)i * BR + r];
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if ((k & (L - 1)) != q) continue;
        const long long o = ((long long)i * BR + r) * NR + k;
        const T dn = c1 * d[o] + c2 * di * (b[o] - acc[r * NR + k]);
        d[o] = dn;
        x_out[o] = x_in[o] + dn;
      }
    }
  }
};

template <int BR, int NR, typename T>
struct MOpDefect {
  static constexpr int kDots = 0;
  const T* b;
  T* res;
  double* partials = nullptr;
  __device__ __forceinline__ void row(int i, const T (&acc)[BR * NR], double*, int q, int L) const {
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if ((k & (L - 1)) != q) continue;
        const long long o = ((long long)i * BR + r) * NR + k;
        res[o] = b[o] - acc[r * NR + k];
      }
  }
};

// ---------------------------------------------------------------- launch
template <int BR, int BC, typename VT, typename XT, int NR, class Op>
static int launch_spmm(const SellMatrix& A, const XT* x, const Op& op, const SpmvTune& t, hipStream_t s) {
  const SellView<VT> v = A.view<VT>();
  if (A.n_slices == 0) return kOk;
  const int xcd = t.xcd_for<VT>();
  const int grid = xcd ? 8 * ((A.n_slices + 7) / 8) : A.n_slices;
  // four waves per slice where the partial accumulators fit in LDS (3-row
  // blocks), one for the 6-row blocks
  if constexpr (BR * NR <= 32) {
    hipLaunchKernelGGL((sell_spmm_kernel<BR, BC, VT, XT, 4, NR, Op>), dim3(grid), dim3(256), 0, s, v, x, op,
                       xcd, t.flags, t.guard);
  } else {
    hipLaunchKernelGGL((sell_spmm_kernel<BR, BC, VT, XT, 1, NR, Op>), dim3(grid), dim3(64), 0, s, v, x, op,
                       xcd, t.flags, t.guard);
  }
  return static_cast<int>(hipGetLastError());
}

#define WN_MR_DISPATCH(A, CALL)                      \
  if ((A).br == 3 && (A).bc == 3) CALL(3, 3)         \
  else if ((A).br == 6 && (A).bc == 6) CALL(6, 6)    \
  else if ((A).br == 3 && (A).bc == 6) CALL(3, 6)    \
  else if ((A).br == 6 && (A).bc == 3) CALL(6, 3)    \
  else return kBadArg;

template <typename VT, typename T, int NR>
static int spmm_store(const SellMatrix& A, const T* x, T* y, int add, const SpmvTune& t, hipStream_t s) {
  using AT = typename AccT<VT>::type;
#define CALL(R, C)                                            \
  {                                                           \
    MOpStore<R, NR, T, AT> op{y, add};                        \
    return launch_spmm<R, C, VT, T, NR>(A, x, op, t, s);      \
  }
  WN_MR_DISPATCH(A, CALL);
#undef CALL
}

template <typename VT, typename T, int NR>
static int spmm_cheb(const SellMatrix& A, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, double c1,
                     double c2, const SpmvTune& t, hipStream_t s) {
#define CALL(R, C)                                                     \
  {                                                                    \
    MOpCheb<R, NR, T> op{b, dinv, x_in, x_out, d, T(c1), T(c2)};       \
    return launch_spmm<R, C, VT, T, NR>(A, x_in, op, t, s);            \
  }
  if (A.br == 3 && A.bc == 3) CALL(3, 3) else if (A.br == 6 && A.bc == 6) CALL(6, 6) else return kBadArg;
#undef CALL
}

template <typename VT, typename T, int NR>
static int spmm_defect(const SellMatrix& A, const T* x, const T* b, T* res, const SpmvTune& t, hipStream_t s) {
#define CALL(R, C)                                            \
  {                                                           \
    MOpDefect<R, NR, T> op{b, res};                           \
    return launch_spmm<R, C, VT, T, NR>(A, x, op, t, s);      \
  }
  if (A.br == 3 && A.bc == 3) CALL(3, 3) else if (A.br == 6 && A.bc == 6) CALL(6, 6) else return kBadArg;
#undef CALL
}

// ---------------------------------------------------------------- multi-column vector kernels
template <typename T, int NR>
__global__ void mr_cheb_first_kernel(long long n, const T* __restrict__ b, const T* __restrict__ dinv,
                                     T* __restrict__ d, T* __restrict__ x, T c2, const int* skip) {
  if (skip && *skip) return;
  const long long total = n * NR;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const T v = c2 * dinv[t / NR] * b[t];
    d[t] = v;
    x[t] = v;
  }
}

// X[i][k] = sum_j inv[i][j] B[j][k] on the coarsest level (one block per row).
template <typename T, int NR>
__global__ __launch_bounds__(256) void mr_coarse_kernel(int n, const T* __restrict__ inv,
                                                        const T* __restrict__ b, T* __restrict__ x,
                                                        const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x;
  double acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double a = double(inv[(size_t)i * n + j]);
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] += a * double(b[(size_t)j * NR + k]);
  }
  __shared__ double sm[4];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const double v = block_sum(acc[k], sm);
    if (threadIdx.x == 0) x[(size_t)i * NR + k] = T(v);
    __syncthreads();
  }
}

static int mr_grid(long long n) {
  long long g = (n + kMrBlock - 1) / kMrBlock;
  if (g > 2048) g = 2048;
  return static_cast<int>(g < 1 ? 1 : g);
}

// Per-level column vectors of the multi-column cycle (b, x, tmp, d, res).
template <typename T>
struct MrLevel {
  T* v[5] = {};
};

struct MultiState {
  int nr = 0;
  long long n = 0;
  const Amg* amg = nullptr;
  int amg_levels = 0;
  std::vector<MrLevel<float>> lv32;
  std::vector<MrLevel<double>> lv64;
  double *r = nullptr, *p = nullptr, *q = nullptr, *zs = nullptr;  // FP64 n x NR
  float *rp32 = nullptr, *zp32 = nullptr;                           // FP32 n x NR
  double* partials = nullptr;  // NR x max blocks
  int max_blocks = 0;
  double* sc = nullptr;  // per-column scalars: kMs* x NR
  int* flags = nullptr;  // [0] all done, [1] iteration, [2..2+NR) done per column, [2+NR..) iterations per column
  double* h_sc = nullptr;
  int* h_flags = nullptr;
  void release() {
    for (auto& l : lv32)
      for (auto* p_ : l.v) if (p_) (void)hipFree(p_);
    for (auto& l : lv64)
      for (auto* p_ : l.v) if (p_) (void)hipFree(p_);
    lv32.clear();
    lv64.clear();
    for (void* p_ : {(void*)r, (void*)p, (void*)q, (void*)zs, (void*)rp32, (void*)zp32, (void*)partials, (void*)sc,
                     (void*)flags})
      if (p_) (void)hipFree(p_);
    if (h_sc) (void)hipHostFree(h_sc);
    if (h_flags) (void)hipHostFree(h_flags);
    r = p = q = zs = nullptr;
    rp32 = zp32 = nullptr;
    partials = sc = h_sc = nullptr;
    flags = h_flags = nullptr;
    n = 0;
    nr = 0;
    amg = nullptr;
  }
};

enum MsIdx { kMsRho = 0, kMsRhoPrev, kMsBeta, kMsAlpha, kMsGamma, kMsRR, kMsRR0, kMsTol2, kMsCount };

static void multi_free(void* p) {
  auto* m = static_cast<MultiState*>(p);
  m->release();
  delete m;
}

static int multi_ensure(Solver& S, long long n, int nr, const Amg* g, int need64cycle, int n_slices_max) {
  if (!S.multi) {
    S.multi = new MultiState();
    S.multi_free = multi_free;
  }
  MultiState& M = *static_cast<MultiState*>(S.multi);
  const int blocks = std::max(n_slices_max + 8, 4096);
  const bool same = M.n >= n && M.nr == nr && M.amg == g && M.max_blocks >= blocks &&
                    (!g || ((int)M.lv32.size() == (int)g->L.size() &&
                            (!need64cycle || (int)M.lv64.size() == (int)g->L.size())));
  if (same && M.r) return kOk;
  M.release();
  M.n = n;
  M.nr = nr;
  M.amg = g;
  M.max_blocks = blocks;
  const size_t nv = (size_t)n * nr;
  WN_HIP(hipMalloc(&M.r, sizeof(double) * nv));
  WN_HIP(hipMalloc(&M.p, sizeof(double) * nv));
  WN_HIP(hipMalloc(&M.q, sizeof(double) * nv));
  WN_HIP(hipMalloc(&M.zs, sizeof(double) * nv));
  WN_HIP(hipMalloc(&M.rp32, sizeof(float) * nv));
  WN_HIP(hipMalloc(&M.zp32, sizeof(float) * nv));
  WN_HIP(hipMalloc(&M.partials, sizeof(double) * nr * (size_t)blocks));
  WN_HIP(hipMalloc(&M.sc, sizeof(double) * kMsCount * nr));
  WN_HIP(hipMalloc(&M.flags, sizeof(int) * (2 + 2 * nr)));
  WN_HIP(hipHostMalloc(&M.h_sc, sizeof(double) * kMsCount * nr, 0));
  WN_HIP(hipHostMalloc(&M.h_flags, sizeof(int) * (2 + 2 * nr), 0));
  if (g) {
    M.lv32.resize(g->L.size());
    if (need64cycle) M.lv64.resize(g->L.size());
    for (size_t l = 0; l < g->L.size(); ++l) {
      const size_t m = (size_t)std::max(g->L[l]->n, 1) * nr;
      for (int k = 0; k < 5; ++k) {
        WN_HIP(hipMalloc(&M.lv32[l].v[k], sizeof(float) * m));
        if (need64cycle) WN_HIP(hipMalloc(&M.lv64[l].v[k], sizeof(double) * m));
      }
    }
  }
  return kOk;
}

template <typename T> static std::vector<MrLevel<T>>& mr_levels(MultiState& M);
template <> std::vector<MrLevel<float>>& mr_levels<float>(MultiState& M) { return M.lv32; }
template <> [[maybe_unused]] std::vector<MrLevel<double>>& mr_levels<double>(MultiState& M) { return M.lv64; }

// The V-cycle of solver.hip on NR columns (same smoother, buffer plan and
// coarse solve).
template <typename VT, typename VTC, typename T, int NR>
static int vcycle_mr(Amg& g, MultiState& M, int l, const T* b, T* x, const int* skip, hipStream_t s) {
  AmgLevel& L = *g.L[l];
  SpmvTune t = g.tune;
  t.flags = skip;
  t.guard = skip ? 1 : 0;
  const long long n = L.n;
  if (l == (int)g.L.size() - 1) {
    const T* inv = coarse_inv<T>(g);
    if (L.n != g.n_coarse || !inv) return kNotReady;
    hipLaunchKernelGGL((mr_coarse_kernel<T, NR>), dim3(std::max(L.n, 1)), dim3(256), 0, s, L.n, inv, b, x, skip);
    return static_cast<int>(hipGetLastError());
  }
  MrLevel<T>& V = mr_levels<T>(M)[l];
  T* tmp = V.v[2];
  T* d = V.v[3];
  T* res = V.v[4];
  const T* dinv = level_dinv<T>(L);
  if (!dinv || !L.has_PR) return kNotReady;
  const int deg = std::max(1, std::min(L.degree, 16));
  const ChebCoef c = cheb_coefs(L.lmax, L.ratio, deg);
  const SellMatrix& A = *L.A;
  int r;
  T* post = (deg % 2 == 0) ? x : tmp;
  T* cur = ((deg - 1) % 2 == 0) ? post : (post == x ? tmp : x);
  T* oth = (cur == x) ? tmp : x;
  hipLaunchKernelGGL((mr_cheb_first_kernel<T, NR>), dim3(mr_grid(n * NR)), dim3(kMrBlock), 0, s, n, b, dinv, d, cur,
                     T(c.c2[0]), skip);
  if ((r = static_cast<int>(hipGetLastError()))) return r;
  for (int k = 1; k < deg; ++k) {
    if ((r = spmm_cheb<VT, T, NR>(A, cur, oth, b, dinv, d, c.c1[k], c.c2[k], t, s))) return r;
    std::swap(cur, oth);
  }
  if ((r = spmm_defect<VT, T, NR>(A, cur, b, res, t, s))) return r;
  MrLevel<T>& N = mr_levels<T>(M)[l + 1];
  if ((r = spmm_store<VT, T, NR>(L.R, res, N.v[0], 0, t, s))) return r;
  if ((r = vcycle_mr<VTC, VTC, T, NR>(g, M, l + 1, N.v[0], N.v[1], skip, s))) return r;
  if ((r = spmm_store<VT, T, NR>(L.P, N.v[1], cur, 1, t, s))) return r;
  oth = (cur == x) ? tmp : x;
  for (int k = 0; k < deg; ++k) {
    if ((r = spmm_cheb<VT, T, NR>(A, cur, oth, b, dinv, d, c.c1[k], c.c2[k], t, s))) return r;
    std::swap(cur, oth);
  }
  return kOk;
}

// ---------------------------------------------------------------- PCG on NR columns
// The vector kernels walk the flat n*NR arrays (fully coalesced); with a
// grid stride that is a multiple of NR every thread stays on one column
// k = t mod NR, and the per-column sums meet in a shuffle over the lanes of
// equal k, then across the waves of the block (fixed order).
template <int NR>
__device__ __forceinline__ void store_column_partials(double v, double* partials) {
  static_assert(NR <= 64 && (NR & (NR - 1)) == 0, "NR must be a power of two <= 64");
#pragma unroll
  for (int off = 32; off >= NR; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double sm[kMrBlock / 64][NR];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane < NR) sm[wid][lane] = v;
  __syncthreads();
  if (threadIdx.x < NR) {
    double t = 0.0;
    for (int w = 0; w < kMrBlock / 64; ++w) t += sm[w][threadIdx.x];
    partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// partial dots of r_k . (S z_k) (z in ZT, S per DOF)
template <typename ZT, int NR>
__global__ __launch_bounds__(kMrBlock) void mr_rho_kernel(long long n, const double* __restrict__ r,
                                                          const ZT* __restrict__ z, const double* __restrict__ scale,
                                                          double* partials, const int* skip) {
  if (skip && skip[0]) return;
  double acc = 0.0;
  const long long total = n * NR;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x)
    acc += r[t] * double(z[t]) * (scale ? scale[t / NR] : 1.0);
  store_column_partials<NR>(acc, partials);
}

// p_k = S z_k + beta_k p_k (beta 0 on the first iteration or for a done column)
template <typename ZT, int NR>
__global__ void mr_update_p_kernel(long long n, const ZT* __restrict__ z, const double* __restrict__ scale,
                                   double* __restrict__ p, const double* __restrict__ sc, const int* flags) {
  if (flags[0]) return;
  const bool first = flags[1] == 0;
  double beta[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) beta[k] = first ? 0.0 : sc[kMsBeta * NR + k];
  const long long total = n * NR;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long i = t / NR;
    const int k = (int)(t - i * NR);
    const double zi = (scale ? scale[i] : 1.0) * double(z[t]);
    p[t] = first ? zi : zi + beta[k] * p[t];
  }
}

// x_k += alpha_k p_k; r_k -= alpha_k q_k; partial ||r_k||^2; rp = S r (FP32)
template <int NR>
__global__ __launch_bounds__(kMrBlock) void mr_update_xr_kernel(long long n, double* __restrict__ x,
                                                                double* __restrict__ r, const double* __restrict__ p,
                                                                const double* __restrict__ q,
                                                                float* __restrict__ rp,
                                                                const double* __restrict__ scale, double* partials,
                                                                const double* __restrict__ sc, const int* flags) {
  if (flags[0]) return;
  const double alpha = sc[kMsAlpha * NR + (threadIdx.x & (NR - 1))];
  double acc = 0.0;
  const long long total = n * NR;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    x[t] += alpha * p[t];
    const double ri = r[t] - alpha * q[t];
    r[t] = ri;
    if (rp) rp[t] = float((scale ? scale[t / NR] : 1.0) * ri);
    acc += ri * ri;
  }
  store_column_partials<NR>(acc, partials);
}

enum MrFin { kMrRho = 0, kMrGamma = 1, kMrRR = 2, kMrStart = 3 };

// One block: column sums of the partials in a fixed order, then the per-column
// scalar updates of the CG recurrences.
template <int NR>
__global__ __launch_bounds__(1024) void mr_finalize_kernel(const double* __restrict__ partials, int nb, int mode,
                                                           double* sc, int* flags, double rtol2) {
  if (mode != kMrStart && flags[0]) return;
  __shared__ double red[NR][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < NR; ++k) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const double* p = partials + (size_t)k * nb;
    const int stride = blockDim.x;
    int i = threadIdx.x;
    for (; i + 3 * stride < nb; i += 4 * stride) {
      a0 += p[i];
      a1 += p[i + stride];
      a2 += p[i + 2 * stride];
      a3 += p[i + 3 * stride];
    }
    for (; i < nb; i += stride) a0 += p[i];
    double s = (a0 + a1) + (a2 + a3);
    s = wave_sum(s);
    if (lane == 0) red[k][wid] = s;
  }
  __syncthreads();
  int* done = flags + 2;
  int* iters = flags + 2 + NR;
  const int k = threadIdx.x;
  if (k < NR) {
  double v = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[k][w];
  switch (mode) {
    case kMrStart:  // v = ||b_k||^2
      sc[kMsRR * NR + k] = v;
      sc[kMsRR0 * NR + k] = v;
      sc[kMsTol2 * NR + k] = rtol2 * v;
      done[k] = (v == 0.0) ? 1 : 0;
      iters[k] = 0;
      break;
    case kMrRho: {
      const double prev = sc[kMsRhoPrev * NR + k];
      sc[kMsRho * NR + k] = v;
      sc[kMsBeta * NR + k] = (done[k] || flags[1] == 0 || prev == 0.0) ? 0.0 : v / prev;
      sc[kMsRhoPrev * NR + k] = v;
      break;
    }
    case kMrGamma:
      sc[kMsGamma * NR + k] = v;
      sc[kMsAlpha * NR + k] = (done[k] || v == 0.0) ? 0.0 : sc[kMsRho * NR + k] / v;
      break;
    case kMrRR:
      if (!done[k]) {
        sc[kMsRR * NR + k] = v;
        iters[k] += 1;
        if (v <= sc[kMsTol2 * NR + k] || !(v == v)) done[k] = 1;
      }
      break;
  }
  }
  __syncthreads();
  if ((mode == kMrRR || mode == kMrStart) && k == 0) {
    int all = 1;
    for (int c = 0; c < NR; ++c) all &= done[c];
    flags[1] = (mode == kMrRR) ? flags[1] + 1 : 0;
    flags[0] = all;
  }
}

template <int NR>
static int mr_fin(MultiState& M, int nb, int mode, double rtol2, hipStream_t s) {
  hipLaunchKernelGGL((mr_finalize_kernel<NR>), dim3(1), dim3(1024), 0, s, M.partials, nb, mode, M.sc, M.flags, rtol2);
  return static_cast<int>(hipGetLastError());
}

// ||b_k||^2 partials and r = b, x = 0, rp = S r (FP32)
template <int NR>
__global__ __launch_bounds__(kMrBlock) void mr_start_kernel(long long n, const double* __restrict__ b,
                                                            double* __restrict__ x, double* __restrict__ r,
                                                            float* __restrict__ rp, const double* __restrict__ scale,
                                                            double* partials) {
  double acc = 0.0;
  const long long total = n * NR;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const double bi = b[t];
    x[t] = 0.0;
    r[t] = bi;
    if (rp) rp[t] = float((scale ? scale[t / NR] : 1.0) * bi);
    acc += bi * bi;
  }
  store_column_partials<NR>(acc, partials);
}

// q = A p with gamma partials (FP64)
template <int NR>
static int mr_spmm_dot(const SellMatrix& A, const double* p, double* q, double* partials, const SpmvTune& t,
                       hipStream_t s) {
#define CALL(R, C)                                                  \
  {                                                                 \
    MOpSpmvDot<R, NR, double> op{q, p, partials};                   \
    return launch_spmm<R, C, double, double, NR>(A, p, op, t, s);   \
  }
  if (A.br == 3 && A.bc == 3) CALL(3, 3) else if (A.br == 6 && A.bc == 6) CALL(6, 6) else return kBadArg;
#undef CALL
}

// grid of the FP64-value SpMM (its block partials feed the CG dots)
static int spmm_grid64(const SellMatrix& A, const SpmvTune& t) {
  return t.xcd_for<double>() ? 8 * ((A.n_slices + 7) / 8) : A.n_slices;
}

// PCG on NR columns of A X = B from X = 0; preconditioner: AMG V-cycle in
// FP32 (vt 1) or FP64 (vt 0) on S A S, z = S M(S r).  B, X: FP64 [n][NR].
template <int NR>
static int pcg_multi_run(Solver& S, const SellMatrix& A, Amg* g, int vt, const double* scale, const double* B,
                         double* X, double rtol, int max_iter, int check_every, const SpmvTune& tune0, hipStream_t s,
                         int* iters, double* rel) {
  const long long n = (long long)A.n_rows * A.br;
  int r;
  if ((r = multi_ensure(S, n, NR, g, vt == 0, A.n_slices))) return r;
  MultiState& M = *static_cast<MultiState*>(S.multi);
  SpmvTune t = tune0;
  t.flags = M.flags;
  t.guard = 1;
  if (check_every <= 0) check_every = 8;
  const int gv = mr_grid(n * NR);
  hipLaunchKernelGGL((mr_start_kernel<NR>), dim3(gv), dim3(kMrBlock), 0, s, n, B, X, M.r, M.rp32, scale, M.partials);
  if ((r = static_cast<int>(hipGetLastError()))) return r;
  if ((r = mr_fin<NR>(M, gv, kMrStart, rtol * rtol, s))) return r;
  const int gs = spmm_grid64(A, t);
  int done_all = 0;
  // chunks sized from each open column's residual reduction rate, as in the
  // single-column PCG (idle iterations after convergence are costly)
  int chunk = check_every, it_prev = 0;
  double rr_prev[NR] = {};
  for (int it = 0; it < max_iter && !done_all;) {
    chunk = std::max(1, std::min(chunk, max_iter - it));
    const int it0 = it;
    for (int c = 0; c < chunk; ++c, ++it) {
      if ((r = vcycle_mr<float, float, float, NR>(*g, M, 0, M.rp32, M.zp32, M.flags, s))) return r;
      hipLaunchKernelGGL((mr_rho_kernel<float, NR>), dim3(gv), dim3(kMrBlock), 0, s, n, M.r, M.zp32, scale,
                         M.partials, M.flags);
      if ((r = static_cast<int>(hipGetLastError()))) return r;
      if ((r = mr_fin<NR>(M, gv, kMrRho, 0.0, s))) return r;
      hipLaunchKernelGGL((mr_update_p_kernel<float, NR>), dim3(mr_grid(n * NR)), dim3(kMrBlock), 0, s, n, M.zp32,
                         scale, M.p, M.sc, M.flags);
      if ((r = static_cast<int>(hipGetLastError()))) return r;
      if ((r = mr_spmm_dot<NR>(A, M.p, M.q, M.partials, t, s))) return r;
      if ((r = mr_fin<NR>(M, gs, kMrGamma, 0.0, s))) return r;
      hipLaunchKernelGGL((mr_update_xr_kernel<NR>), dim3(gv), dim3(kMrBlock), 0, s, n, X, M.r, M.p, M.q, M.rp32,
                         scale, M.partials, M.sc, M.flags);
      if ((r = static_cast<int>(hipGetLastError()))) return r;
      if ((r = mr_fin<NR>(M, gv, kMrRR, 0.0, s))) return r;
    }
    WN_HIP(hipMemcpyAsync(M.h_flags, M.flags, sizeof(int) * (2 + 2 * NR), hipMemcpyDeviceToHost, s));
    WN_HIP(hipMemcpyAsync(M.h_sc, M.sc, sizeof(double) * kMsCount * NR, hipMemcpyDeviceToHost, s));
    WN_HIP(hipStreamSynchronize(s));
    done_all = M.h_flags[0];
    if (done_all) break;
    double need_max = 0.0;
    bool known = true;
    for (int k = 0; k < NR; ++k) {
      if (M.h_flags[2 + k]) continue;
      const double rr = M.h_sc[kMsRR * NR + k], tol2 = M.h_sc[kMsTol2 * NR + k];
      const double prev = it_prev == 0 ? M.h_sc[kMsRR0 * NR + k] : rr_prev[k];
      rr_prev[k] = rr;
      if (!(rr > 0.0 && rr < prev && tol2 > 0.0 && tol2 < rr)) {
        known = false;
        continue;
      }
      need_max = std::max(need_max, std::log(tol2 / rr) / (std::log(rr / prev) / (it - it0)));
    }
    it_prev = it;
    chunk = (known && need_max < check_every) ? std::max(1, (int)need_max) : check_every;
  }
  WN_HIP(hipMemcpyAsync(M.h_sc, M.sc, sizeof(double) * kMsCount * NR, hipMemcpyDeviceToHost, s));
  WN_HIP(hipMemcpyAsync(M.h_flags, M.flags, sizeof(int) * (2 + 2 * NR), hipMemcpyDeviceToHost, s));
  WN_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < NR; ++k) {
    if (iters) iters[k] = M.h_flags[2 + NR + k];
    const double rr0 = M.h_sc[kMsRR0 * NR + k];
    if (rel) rel[k] = rr0 > 0.0 ? std::sqrt(M.h_sc[kMsRR * NR + k] / rr0) : 0.0;
  }
  return kOk;
}

int pcg_multi(Solver& S, const SellMatrix& A, Amg* g, int vt, const double* scale, int nr, const double* B, double* X,
              double rtol, int max_iter, int check_every, const SpmvTune& t, hipStream_t s, int* iters, double* rel) {
  if (!g || vt != 1 || (g->store & 2) == 0) return kBadMode;
  switch (nr) {
    case 4: return pcg_multi_run<4>(S, A, g, vt, scale, B, X, rtol, max_iter, check_every, t, s, iters, rel);
    case 8: return pcg_multi_run<8>(S, A, g, vt, scale, B, X, rtol, max_iter, check_every, t, s, iters, rel);
    default: return kBadArg;
  }
}

// Y = A X on nr interleaved columns (prec 0 FP64, 1 FP32)
int bsr_spmm(const SellMatrix& A, int prec, int nr, const void* X, void* Y, const SpmvTune& t, hipStream_t s) {
  SpmvTune tt = t;
  tt.flags = nullptr;
  tt.guard = 0;
  if (prec == 0) {
    if (!A.d_v64) return kNotReady;
    if (nr == 4) return spmm_store<double, double, 4>(A, (const double*)X, (double*)Y, 0, tt, s);
    if (nr == 8) return spmm_store<double, double, 8>(A, (const double*)X, (double*)Y, 0, tt, s);
    return kBadArg;
  }
  if (prec == 1) {
    if (!A.d_v32) return kNotReady;
    if (nr == 4) return spmm_store<float, float, 4>(A, (const float*)X, (float*)Y, 0, tt, s);
    if (nr == 8) return spmm_store<float, float, 8>(A, (const float*)X, (float*)Y, 0, tt, s);
    return kBadArg;
  }
  return kBadMode;
}

}  // namespace wn
