// Smoothed-aggregation AMG V-cycle and the PCG / iterative-refinement drivers.
//
// V-cycle (per level l < coarsest), stored values VT (double, float or
// _Float16) with vectors and arithmetic T (double for VT = double, float
// otherwise):
//   pre-smooth  : Chebyshev(deg) on D^-1 A from x = 0   (1 elementwise + deg-1 SpMV kernels)
//   defect      : res = b - A x                        (1 SpMV)
//   restrict    : b_{l+1} = R res                      (1 SpMV with R = P^T stored explicitly)
//   recurse
//   prolongate  : x += P x_{l+1}                       (1 SpMV)
//   post-smooth : Chebyshev(deg) from x                (deg SpMV kernels)
// Coarsest level: x = A_c^{-1} b with the exact inverse (computed in FP64,
// stored as packed upper-triangle tiles in the cycle's type, products
// accumulated in FP64).  Pre- and post-smoothers use the same polynomial, so
// the cycle is a symmetric preconditioner for PCG.
//
// PCG keeps rho/alpha/beta/||r||^2 on the device; a done flag raised by the
// x/r update makes every later kernel of the loop return at once, so the host
// enqueues iterations in chunks and checks the flag only between chunks.  The
// preconditioner may run in a lower precision than the Krylov vectors and on
// a diagonally scaled operator: z = S M(S r); the precision conversions and
// the scaling are fused into the rho, p-update and x/r-update kernels.
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <utility>
#include "solver.hpp"

namespace wn {

int amg_alloc_vectors(Amg& g, int prec_mask) {
  for (auto& lp : g.L) {
    AmgLevel& l = *lp;
    const size_t n = std::max(l.n, 1);
    for (int k = 0; k < 5; ++k) {
      if ((prec_mask & 1) && !l.v64[k]) WN_HIP(hipMalloc(&l.v64[k], sizeof(double) * n));
      if ((prec_mask & 6) && !l.v32[k]) WN_HIP(hipMalloc(&l.v32[k], sizeof(float) * n));
    }
  }
  return kOk;
}

// d = c2 * dinv .* b;  x = d
template <typename T>
__global__ void cheb_first_kernel(long long n, const T* __restrict__ b, const T* __restrict__ dinv, T* __restrict__ d,
                                  T* __restrict__ x, T c2, const int* skip) {
  if (skip && *skip) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const T v = c2 * dinv[i] * b[i];
    d[i] = v;
    x[i] = v;
  }
}

// x = inv * b on the coarsest level: one 256-thread block per row; the
// inverse is stored in the cycle's type T (its bytes dominate), the dot
// product accumulates in FP64.
template <typename T>
__global__ __launch_bounds__(256) void coarse_gemv_kernel(int n, const T* __restrict__ inv,
                                                          const T* __restrict__ b, T* __restrict__ x,
                                                          const int* skip) {
  if (skip && *skip) return;
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) s += double(inv[(size_t)i * n + j]) * double(b[j]);
  __shared__ double sm[4];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) x[i] = T(s);
}

// ---- packed symmetric coarse inverse ----------------------------------------
// The inverse is symmetric, so only its upper-triangle 64 x 64 tiles (I <= J)
// are kept: tile k = I*nt - I(I-1)/2 + (J - I), tile-major, each tile
// row-major, zero beyond n.  One 256-thread block per tile reads the tile once
// and produces both products it takes part in: the row partial A_IJ x_J (for
// y_I) and, off the diagonal, the column partial A_IJ^T x_I (for y_J).  The
// partials (FP64) go to part[(2k + side) * 64 + r]; a second kernel sums, for
// each tile row, its nt partials in a fixed order, so the result is
// reproducible bit for bit.  This halves the bytes of the dense product.
constexpr int kTile = 64;

__host__ __device__ inline long long sym_tile_index(int I, int J, int nt) {
  return (long long)I * nt - (long long)I * (I - 1) / 2 + (J - I);
}

template <typename T>
__global__ __launch_bounds__(256) void sym_tile_pack_kernel(int n, int nt, const double* __restrict__ inv,
                                                           const int2* __restrict__ tij, T* __restrict__ tiles) {
  const long long k = blockIdx.x;
  const int2 ij = tij[k];
  T* dst = tiles + k * (kTile * kTile);
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int i = ij.x * kTile + e / kTile, j = ij.y * kTile + e % kTile;
    dst[e] = (i < n && j < n) ? T(inv[(size_t)i * n + j]) : T(0);
  }
}

// v from another lane of its 16-lane DPP row (both 32-bit halves)
template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(u & 0xffffffffu), Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), Ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// Sum over the 16 lanes of a DPP row; every lane of the row gets the same
// bits (each step adds two equal-bit partners up to commutation).
__device__ __forceinline__ double dpp_row16_sum(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141>(v);  // row_half_mirror: lane i <- 7 - i within 8
  v += dpp_f64<0x140>(v);  // row_mirror: lane i <- 15 - i within 16
  return v;
}

// x[i] for i < n, else 0 (the tile's padding rows and columns): the load
// reads an in-bounds element either way, so that no branch around it makes
// the wavefront wait for every load in flight before the next one (n >= 1)
template <typename XT>
__device__ __forceinline__ double x_or_zero(const XT* __restrict__ x, int i, int n) {
  const double v = double(x[i < n ? i : n - 1]);
  return i < n ? v : 0.0;
}

// Thread t of the block holds V = 16 B / sizeof(T) consecutive columns
// col0 = (t % CPR) * V of NI rows t / CPR + RPI * m: every load instruction
// of the block covers 4 KiB of the tile contiguously.  T: tile values, XT:
// the vector (an FP32 cycle may apply the FP64 inverse).
template <typename T, typename XT = T>
__global__ __launch_bounds__(256) void sym_tile_gemv_kernel(int n, const T* __restrict__ tiles,
                                                           const int2* __restrict__ tij, const XT* __restrict__ x,
                                                           double* __restrict__ part, const int* skip) {
  if (skip && *skip) return;
  constexpr int V = 16 / sizeof(T);
  constexpr int CPR = kTile / V;    // 16-byte chunks per tile row
  constexpr int RPI = 256 / CPR;    // rows per block-wide load instruction
  constexpr int NI = kTile / RPI;   // load instructions per thread
  using Chunk = VecN<T, V>;
  const long long k = blockIdx.x;
  const int2 ij = tij[k];
  const int t = threadIdx.x;
  const int c0 = (t % CPR) * V, r0 = t / CPR;
  const Chunk* src = reinterpret_cast<const Chunk*>(tiles + k * (kTile * kTile));
  // non-temporal: the 200-400 MB of tiles streamed once per cycle would
  // otherwise displace the cycle's and the CG iteration's reused vectors
  // from L2 and the Infinity Cache (I-beam: CG vector work 52 -> 37 us per
  // iteration, mixed solve 20.6 -> 19.8 ms)
  Chunk a[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m) load_vec<T, V>(reinterpret_cast<const T*>(src + t + 256 * m), a[m].e);
  double xj[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int j = ij.y * kTile + c0 + v;
    xj[v] = x_or_zero(x, j, n);
  }
  double rowp[NI], colp[V];
#pragma unroll
  for (int v = 0; v < V; ++v) colp[v] = 0.0;
#pragma unroll
  for (int m = 0; m < NI; ++m) {
    const int i = ij.x * kTile + r0 + RPI * m;
    const double xi = x_or_zero(x, i, n);
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double av = double(a[m].e[v]);
      s += av * xj[v];
      colp[v] += av * xi;
    }
    rowp[m] = s;
  }
  // row partials: sum over the CPR lanes of a row, 16 of them in one DPP row
  // (quad swaps, then the half-row and row mirrors), plus one shuffle when a
  // row spans 32 lanes
#pragma unroll
  for (int m = 0; m < NI; ++m) {
    rowp[m] = dpp_row16_sum(rowp[m]);
    if (CPR == 32) rowp[m] += __shfl_xor(rowp[m], 16, 64);
  }
  double* prow = part + (2 * k) * kTile;
  if (t % CPR == 0) {
#pragma unroll
    for (int m = 0; m < NI; ++m) prow[r0 + RPI * m] = rowp[m];
  }
  if (ij.x == ij.y) return;  // the diagonal tile has no transposed product
  // column partials: every lane's V column sums meet in LDS, one thread per
  // column adds the block's RPI row groups in order
  __shared__ double cs[256 / (kTile / 4)][kTile];
#pragma unroll
  for (int v = 0; v < V; ++v) cs[r0][c0 + v] = colp[v];
  __syncthreads();
  if (t < kTile) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RPI; ++q) a += cs[q][t];
    part[(2 * k + 1) * kTile + t] = a;
  }
}

// y_I = sum over m of (m < I ? column partial of tile (m, I) : row partial of
// tile (I, m)); kGroups groups of 64 threads take every kGroups-th m, then
// add in group order.
constexpr int kRedGroups = 16;

template <typename T>
__global__ __launch_bounds__(kTile * kRedGroups) void sym_tile_reduce_kernel(int n, int nt,
                                                                            const double* __restrict__ part,
                                                                            T* __restrict__ y, const int* skip) {
  if (skip && *skip) return;
  const int I = blockIdx.x;
  const int g = threadIdx.x / kTile, r = threadIdx.x % kTile;
  double s = 0.0;
  for (int m = g; m < nt; m += kRedGroups) {
    const long long idx = m < I ? 2 * sym_tile_index(m, I, nt) + 1 : 2 * sym_tile_index(I, m, nt);
    s += part[idx * kTile + r];
  }
  __shared__ double red[kRedGroups][kTile];
  red[g][r] = s;
  __syncthreads();
  const int i = I * kTile + threadIdx.x;
  if (threadIdx.x < kTile && i < n) {
    double v = red[0][threadIdx.x];
#pragma unroll
    for (int k = 1; k < kRedGroups; ++k) v += red[k][threadIdx.x];
    y[i] = T(v);
  }
}

template <typename T>
static int coarse_apply(const Amg& g, const T* b, T* x, const int* skip, hipStream_t s) {
  // an FP32 cycle with coarse_fp64 set applies the FP64 inverse (FP64
  // products and partial sums; only the vectors stay FP32)
  if (std::is_same<T, float>::value && g.coarse_fp64 && g.tiles64 && g.tpart && g.tij) {
    hipLaunchKernelGGL((sym_tile_gemv_kernel<double, T>), dim3((unsigned)g.ntiles), dim3(256), 0, s, g.n_coarse,
                       g.tiles64, g.tij, b, g.tpart, skip);
    hipLaunchKernelGGL(sym_tile_reduce_kernel<T>, dim3(g.nt), dim3(kTile * kRedGroups), 0, s, g.n_coarse, g.nt,
                       g.tpart, x, skip);
    return static_cast<int>(hipGetLastError());
  }
  const T* tiles = coarse_tiles<T>(g);
  if (tiles && g.tpart && g.tij) {
    hipLaunchKernelGGL(sym_tile_gemv_kernel<T>, dim3((unsigned)g.ntiles), dim3(256), 0, s, g.n_coarse, tiles, g.tij,
                       b, g.tpart, skip);
    hipLaunchKernelGGL(sym_tile_reduce_kernel<T>, dim3(g.nt), dim3(kTile * kRedGroups), 0, s, g.n_coarse, g.nt,
                       g.tpart, x, skip);
    return static_cast<int>(hipGetLastError());
  }
  const T* inv = coarse_inv<T>(g);
  if (!inv) return kNotReady;
  hipLaunchKernelGGL(coarse_gemv_kernel<T>, dim3(std::max(g.n_coarse, 1)), dim3(256), 0, s, g.n_coarse, inv, b, x,
                     skip);
  return static_cast<int>(hipGetLastError());
}

// One V-cycle from level l.  VT: values of this level; VTC: values of every
// coarser level; T: vector and arithmetic type; VS / VSC: values of the
// smoothing steps on this level / the coarser ones (the defect, restriction
// and prolongation use VT / VTC).  With VS != VT the cycle is not symmetric:
// it serves flexible PCG.
// The buffer a level's pre-smoother starts in (its first iterate): the
// post-smoother (deg steps, ping-pong) must end in x, so it starts in
// `post`; the pre-smoother (1 elementwise + deg-1 product steps) is started
// so that it ends in `post` too, and the prolongation adds in place there.
template <typename T>
static T* first_iterate(int deg, T* x, T* tmp) {
  T* post = (deg % 2 == 0) ? x : tmp;
  return ((deg - 1) % 2 == 0) ? post : (post == x ? tmp : x);
}
static int level_degree(const AmgLevel& L) { return std::max(1, std::min(L.degree, 16)); }

// first_done: the caller's restriction wrote this level's first Chebyshev
// iterate already (see spmv_restrict_first).
template <typename VT, typename VTC, typename T, typename VS = VT, typename VSC = VTC>
static int vcycle(Amg& g, int l, const T* b, T* x, const int* skip, hipStream_t s, const RhoReq* rq = nullptr,
                  bool first_done = false) {
  AmgLevel& L = *g.L[l];
  SpmvTune t = g.tune;
  t.flags = skip;
  t.guard = skip ? 1 : 0;
  const long long n = L.n;
  if (l == (int)g.L.size() - 1) {
    if (L.n != g.n_coarse) return kNotReady;
    return coarse_apply<T>(g, b, x, skip, s);
  }
  T** v = level_vecs<T>(L);
  T* tmp = v[2];
  T* d = v[3];
  T* res = v[4];
  const T* dinv = level_dinv<T>(L);
  if (!tmp || !dinv || !L.has_PR) return kNotReady;
  const int deg = level_degree(L);
  const ChebCoef c = cheb_coefs(L.lmax, L.ratio, deg);
  const SellMatrix& A = *L.A;
  int r;
  T* cur = first_iterate(deg, x, tmp);  // no copies: see first_iterate
  T* oth = (cur == x) ? tmp : x;
  // the finest level of an FP32 / FP64 cycle may run its products matrix-free
  const bool mf = l == 0 && std::is_same<VT, T>::value && mf_smooth_ready<T>(g.mfs);
  // rq: the last post-smoothing step (matrix-free, patch layout) also writes
  // the partials of rho = r . (S x) for the Krylov loop
  if (rq && !mf) return kBadMode;
  // a unit inverse diagonal is passed to the matrix-free steps as null
  const T* dinv_mf = level_unit<T>(L) ? nullptr : dinv;
  auto cheb = [&](const T* xi, T* xo, int k, const RhoReq* q = nullptr) -> int {
    if (mf && q) return mf_smooth_cheb_rho<T>(g.mfs, xi, xo, b, dinv_mf, d, T(c.c1[k]), T(c.c2[k]), *q, skip, s);
    if (mf) return mf_smooth_cheb<T>(g.mfs, xi, xo, b, dinv_mf, d, T(c.c1[k]), T(c.c2[k]), skip, s);
    return spmv_cheb<VS, T>(A, xi, xo, b, dinv, d, c.c1[k], c.c2[k], t, s);
  };
  if (!first_done) {
    hipLaunchKernelGGL(cheb_first_kernel<T>, dim3(vec_grid(n)), dim3(256), 0, s, n, b, dinv, d, cur, T(c.c2[0]),
                       skip);
    if ((r = static_cast<int>(hipGetLastError()))) return r;
  }
  for (int k = 1; k < deg; ++k) {
    if ((r = cheb(cur, oth, k))) return r;
    std::swap(cur, oth);
  }
  // defect and restriction
  if ((r = mf ? mf_smooth_defect<T>(g.mfs, cur, b, res, skip, s) : spmv_defect<VT, T>(A, cur, b, res, t, s))) return r;
  AmgLevel& N = *g.L[l + 1];
  T** nv = level_vecs<T>(N);
  // a smoothed coarser level gets its first Chebyshev iterate from the
  // restriction's epilogue (one launch fewer per cycle)
  const T* dinv_n = level_dinv<T>(N);
  const bool fuse_first = l + 1 < (int)g.L.size() - 1 && nv[1] && nv[2] && nv[3] && dinv_n;
  if (fuse_first) {
    const int deg_n = level_degree(N);
    const ChebCoef cn = cheb_coefs(N.lmax, N.ratio, deg_n);
    if ((r = spmv_restrict_first<VT, T>(L.R, res, nv[0], dinv_n, first_iterate(deg_n, nv[1], nv[2]), nv[3],
                                        cn.c2[0], t, s)))
      return r;
  } else if ((r = spmv_restrict<VT, T>(L.R, res, nv[0], t, s))) {
    return r;
  }
  if ((r = vcycle<VTC, VTC, T, VSC, VSC>(g, l + 1, nv[0], nv[1], skip, s, nullptr, fuse_first))) return r;
  if ((r = spmv_addto<VT, T>(L.P, nv[1], cur, t, s))) return r;  // cur == post here
  oth = (cur == x) ? tmp : x;
  for (int k = 0; k < deg; ++k) {
    if ((r = cheb(cur, oth, k, k == deg - 1 ? rq : nullptr))) return r;
    std::swap(cur, oth);
  }
  return kOk;
}

template <typename T>
bool amg_first_iter(Amg& g, int vt, T* z, FirstIter<T>& f) {
  const bool kind = std::is_same<T, double>::value ? (vt == 0 && (g.store & 1)) : (vt == 1 && (g.store & 2));
  if (!kind || g.L.size() < 2 || !z) return false;
  AmgLevel& L = *g.L[0];
  T** v = level_vecs<T>(L);
  const T* dinv = level_dinv<T>(L);
  if (!v[2] || !v[3] || !dinv || !L.has_PR) return false;
  const int deg = level_degree(L);
  const ChebCoef c = cheb_coefs(L.lmax, L.ratio, deg);
  f.dinv = dinv;
  f.unit = level_unit<T>(L);
  f.x1 = first_iterate(deg, z, v[2]);
  f.d1 = v[3];
  f.c2 = T(c.c2[0]);
  return true;
}
template bool amg_first_iter<double>(Amg&, int, double*, FirstIter<double>&);
template bool amg_first_iter<float>(Amg&, int, float*, FirstIter<float>&);

template <typename T>
int amg_first_launch(const FirstIter<T>& f, long long n, const T* b, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(cheb_first_kernel<T>, dim3(vec_grid(n)), dim3(256), 0, s, n, b, f.dinv, f.d1, f.x1, f.c2, skip);
  return static_cast<int>(hipGetLastError());
}
template int amg_first_launch<double>(const FirstIter<double>&, long long, const double*, const int*, hipStream_t);
template int amg_first_launch<float>(const FirstIter<float>&, long long, const float*, const int*, hipStream_t);

// Partials the cycle of kind vt writes for a RhoReq (0: that cycle cannot
// form rho: its finest level is not smoothed matrix-free in the patch layout,
// or vt is not 0 or 1).
int amg_rho_blocks(const Amg& g, int vt) {
  if (g.L.size() < 2) return 0;
  if (vt == 0) return (g.store & 1) ? mf_smooth_rho_blocks<double>(g.mfs) : 0;
  if (vt == 1) return (g.store & 2) ? mf_smooth_rho_blocks<float>(g.mfs) : 0;
  return 0;
}

// One V-cycle, z = M r; with rq also the partials of rho = r . (S z) (see
// amg_rho_blocks).  vt: 0 FP64 values and vectors, 1 FP32 values, 2 FP16
// values on every level, 3 FP16 values on the finest level and FP32 below it,
// 4 FP32 on the finest level and FP16 below it (vt >= 1: FP32 vectors and
// arithmetic).  (FP16 values in the finest smoother with an FP32 defect make
// the cycle nonsymmetric by O(eps16) and CG then stalls, so that split is not
// offered.)
int amg_apply(Amg& g, int vt, const void* r, void* z, const int* skip, hipStream_t s, const RhoReq* rq,
              bool first_done) {
  if (g.L.empty()) return kNotReady;
  const float* r32 = static_cast<const float*>(r);
  float* z32 = static_cast<float*>(z);
  if (rq && amg_rho_blocks(g, vt) <= 0) return kBadMode;
  switch (vt) {
    case 0:
      if (!(g.store & 1)) return kNotReady;
      return vcycle<double, double, double>(g, 0, static_cast<const double*>(r), static_cast<double*>(z), skip, s,
                                            rq, first_done);
    case 1:
      if (!(g.store & 2)) return kNotReady;
      return vcycle<float, float, float>(g, 0, r32, z32, skip, s, rq, first_done);
    case 2:
      if (!(g.store & 4)) return kNotReady;
      return vcycle<_Float16, _Float16, float>(g, 0, r32, z32, skip, s);
    case 3:
      if ((g.store & 6) != 6) return kNotReady;
      return vcycle<_Float16, float, float>(g, 0, r32, z32, skip, s);
    case 4:
      if ((g.store & 6) != 6) return kNotReady;
      return vcycle<float, _Float16, float>(g, 0, r32, z32, skip, s);
    case 5:  // FP16 smoothing steps, FP32 defects and transfers: nonsymmetric
      if ((g.store & 6) != 6) return kNotReady;
      return vcycle<float, float, float, _Float16, _Float16>(g, 0, r32, z32, skip, s);

    default:
      return kBadMode;
  }
}

int amg_set_coarse_fp64(Amg& g, int fp64) {
  if (fp64 && !g.tiles64) return kNotReady;
  g.coarse_fp64 = fp64 ? 1 : 0;
  return kOk;
}

// `reps` Chebyshev steps (the k = 1 coefficients) of level l < last on the
// level's own vectors (b, x <-> tmp, d), with the level's assembled values in
// the precision of vt (0 FP64, 1 FP32): the smoother kernel of a V-cycle on
// its own, so that its time per launch can be measured.  The vectors hold
// whatever the last cycle left in them.
int amg_level_cheb(Amg& g, int l, int vt, int reps, hipStream_t s) {
  if (l < 0 || l + 1 >= (int)g.L.size() || reps < 0) return kBadArg;
  AmgLevel& L = *g.L[l];
  const int deg = std::max(2, std::min(L.degree, 16));
  const ChebCoef c = cheb_coefs(L.lmax, L.ratio, deg);
  SpmvTune t = g.tune;
  t.flags = nullptr;
  t.guard = 0;
  auto run = [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    T** v = level_vecs<T>(L);
    const T* dinv = level_dinv<T>(L);
    if (!v[0] || !v[1] || !v[2] || !v[3] || !dinv) return kNotReady;
    T* cur = v[1];
    T* oth = v[2];
    for (int i = 0; i < reps; ++i) {
      const int r = spmv_cheb<T, T>(*L.A, cur, oth, v[0], dinv, v[3], c.c1[1], c.c2[1], t, s);
      if (r) return r;
      std::swap(cur, oth);
    }
    return kOk;
  };
  if (vt == 0) return (g.store & 1) ? run((double*)nullptr) : kNotReady;
  if (vt == 1) return (g.store & 2) ? run((float*)nullptr) : kNotReady;
  return kBadMode;
}

// ------------------------------------------------------------------ PCG
__global__ void pcg_start_kernel(double* sc, int* flags, double rtol2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc[kScRR0] = sc[kScRR];
    sc[kScTol2] = rtol2 * sc[kScRR];
    flags[0] = (sc[kScRR] == 0.0) ? 1 : 0;
    flags[1] = 0;
  }
}

// ------------------------------------------------------------------ graph keys
template <typename P>
static void key_push(std::vector<uint64_t>& k, P v) {
  uint64_t u = 0;
  static_assert(sizeof(P) <= sizeof(u), "key item too wide");
  std::memcpy(&u, &v, sizeof(P));
  k.push_back(u);
}

static void key_matrix(std::vector<uint64_t>& k, const SellMatrix& A) {
  for (int v : {A.br, A.bc, A.n_rows, A.n_cols, A.n_slices, A.sigma, A.lanes}) key_push(k, v);
  key_push(k, A.slots);
  for (const void* p : {(const void*)A.d_off, (const void*)A.d_rinfo, (const void*)A.d_col,
                        (const void*)A.d_v64, (const void*)A.d_v64e, (const void*)A.d_v32, (const void*)A.d_v16})
    key_push(k, p);
}

static void key_tune(std::vector<uint64_t>& k, const SpmvTune& t) {
  for (int v : {t.waves, t.unroll, t.xcd_remap, t.waves64, t.unroll64, t.xcd_remap64, t.guard}) key_push(k, v);
  key_push(k, t.flags);
}

// every pointer and size the matrix-free kernels take (both layouts)
static void key_mf(std::vector<uint64_t>& k, const MfOp& m) {
  for (const void* p : {(const void*)m.qd, (const void*)m.fe, (const void*)m.mask, (const void*)m.conn,
                        (const void*)m.fix, (const void*)m.E, (const void*)m.nu, (const void*)m.n2e,
                        (const void*)m.n2e_ptr, (const void*)m.chunk, (const void*)m.pe, (const void*)m.pnode,
                        (const void*)m.pconn, (const void*)m.poff, (const void*)m.pslot, (const void*)m.nhead,
                        (const void*)m.novf, (const void*)m.qdp, (const void*)m.part})
    key_push(k, p);
  for (int v : {m.M, m.n_nodes, m.n_chunks, m.affine, m.patch, m.n_patch, m.stride, m.n_local, m.gen}) key_push(k, v);
}

static void key_amg(std::vector<uint64_t>& k, const Amg* g) {
  key_push(k, g);
  if (!g) return;
  for (int v : {g->store, g->n_coarse, (int)g->L.size()}) key_push(k, v);
  for (const void* p : {(const void*)g->inv, (const void*)g->inv32, (const void*)g->tiles64, (const void*)g->tiles32,
                        (const void*)g->tij, (const void*)g->tpart, (const void*)g->mfs.op,
                        (const void*)g->mfs.s64, (const void*)g->mfs.s64e, (const void*)g->mfs.s32,
                        (const void*)g->mfs.s32e, (const void*)g->mfs.qd32, (const void*)g->mfs.fe32,
                        (const void*)g->mfs.qdp32, (const void*)g->mfs.part32})
    key_push(k, p);
  key_push(k, g->mfs.gen);
  if (g->mfs.op) key_mf(k, *g->mfs.op);
  key_tune(k, g->tune);
  for (const auto& lp : g->L) {
    const AmgLevel& l = *lp;
    for (int v : {l.n, l.degree, (int)l.has_PR}) key_push(k, v);
    key_push(k, l.lmax);
    key_push(k, l.ratio);
    key_push(k, l.dinv64);
    key_push(k, l.dinv32);
    for (auto* p : l.v64) key_push(k, p);
    for (auto* p : l.v32) key_push(k, p);
    if (l.A) key_matrix(k, *l.A);
    if (l.has_PR) {
      key_matrix(k, l.P);
      key_matrix(k, l.R);
    }
  }
}

// Capture `enqueue(stream)` into S.gexec unless a graph with this key exists.
template <typename F>
static int graph_for(Solver& S, const std::vector<uint64_t>& key, F&& enqueue) {
  if (S.gexec && S.gkey == key) return kOk;
  S.drop_graph();
  if (!S.cap) WN_HIP(hipStreamCreateWithFlags(&S.cap, hipStreamNonBlocking));
  WN_HIP(hipStreamBeginCapture(S.cap, hipStreamCaptureModeThreadLocal));
  const int r = enqueue(S.cap);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(S.cap, &graph);
  if (r || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return r ? r : static_cast<int>(e);
  }
  const hipError_t ei = hipGraphInstantiate(&S.gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    S.gexec = nullptr;
    return static_cast<int>(ei);
  }
  S.gkey = key;
  return kOk;
}

// PCG on A x = b from x = 0, vectors in T.  Preconditioner kinds: 0 none,
// 1 Jacobi (dinv in T), 2 AMG (amg_vt as amg_apply; scale S optional: the
// preconditioner is z = S M(S r)).
template <typename T>
int pcg_run(Solver& S, const SellMatrix& A, int pkind, const T* dinv, Amg* amg, int amg_vt, const double* scale,
            const T* b, T* x, double rtol, int max_iter, int check_every, const SpmvTune& tune0, hipStream_t s,
            int* iters, double* rel, const MfOp* mf = nullptr) {
  KVec<T>& k = *reinterpret_cast<KVec<T>*>(std::is_same<T, double>::value ? (void*)&S.k64 : (void*)&S.k32);
  // with mf the operator is matrix-free (FP64 vectors only) and A is unused
  if (mf && !std::is_same<T, double>::value) return kBadMode;
  const long long n = mf ? 3LL * mf->n_nodes : (long long)A.n_rows * A.br;
  int r;
  if ((r = S.init())) return r;
  if ((r = k.ensure(n))) return r;
  // a nonsymmetric cycle needs the flexible (Polak-Ribiere) beta, which reads
  // q = A p of the previous iteration in the rho kernel
  const bool flexible = pkind == 2 && amg_vt == 5;
  // rho from the cycle's last smoothing step (FP64 Krylov vectors, a
  // symmetric cycle whose finest level is smoothed matrix-free in the patch
  // layout): that step writes the rho partials, the rho pass is skipped
  const int nb_cycle_rho =
      (pkind == 2 && amg && !flexible && std::is_same<T, double>::value) ? amg_rho_blocks(*amg, amg_vt) : 0;
  if ((r = S.ws.alloc(std::max((mf ? mf_grid(*mf) : A.n_slices), nb_cycle_rho) + 8))) return r;
  SpmvTune t = tune0;
  t.flags = S.flags;
  t.guard = 1;
  if (check_every <= 0) check_every = 8;
  // precision of the preconditioner vectors; a separate FP32 pair is needed
  // when the V-cycle runs in FP32 under FP64 vectors, or when S is applied
  const bool amg32 = (pkind == 2) && amg_vt != 0;
  if (pkind == 2 && !amg32 && !std::is_same<T, double>::value) return kBadMode;
  const bool sep32 = amg32 && (!std::is_same<T, float>::value || scale != nullptr);
  const bool sep64 = (pkind == 2) && !amg32 && scale != nullptr;
  if (sep32 && (r = S.ensure_p(n))) return r;
  float* rp32 = sep32 ? S.rp32 : nullptr;
  // x = 0, r = b, rr = b.b
  if ((r = launch_fill<T>(n, x, T(0), s))) return r;
  if (flexible && (r = launch_fill<T>(n, k.q, T(0), s))) return r;
  if ((r = launch_cast_scale<T, T>(n, b, nullptr, 1.0, k.r, s))) return r;
  if ((r = launch_dot<T>(n, k.r, k.r, S.ws, S.ws.sc + kScRR, s))) return r;
  hipLaunchKernelGGL(pcg_start_kernel, dim3(1), dim3(64), 0, s, S.ws.sc, S.flags, rtol * rtol);
  if (sep32 && (r = launch_cast_scale<T, float>(n, k.r, scale, 1.0, rp32, s))) return r;
  if (sep64 && (r = launch_cast_scale<T, T>(n, k.r, scale, 1.0, k.z, s))) return r;
  // The finest level's first Chebyshev iterate of each cycle is formed by
  // the kernel that writes the cycle's input, S r: the x/r update of the
  // previous iteration, and one launch here for the first iteration.
  FirstIter<float> fi32;
  FirstIter<T> fi64;
  bool first_done = false;
  if (pkind == 2 && !flexible && sep32) {
    first_done = amg_first_iter<float>(*amg, amg_vt, S.zp32, fi32) &&
                 !(r = amg_first_launch<float>(fi32, n, rp32, nullptr, s));
  } else if (pkind == 2 && sep64 && std::is_same<T, double>::value) {
    first_done = amg_first_iter<T>(*amg, 0, k.q, fi64) && !(r = amg_first_launch<T>(fi64, n, k.z, nullptr, s));
  }
  if (r) return r;
  // One iteration; the first-iteration switch is read on the device from the
  // iteration counter, so every iteration enqueues the same kernels.
  // Unless the cycle is flexible, rho and gamma = p . A p are finalised by
  // the kernels that use them (every block of the p update sums the rho
  // partials, every block of the x, r update the gamma partials), so an
  // iteration launches one one-block finalisation (rr and the done flag)
  // instead of three.
  const bool fused = !flexible;
  const RhoReq rq{reinterpret_cast<const double*>(k.r), scale, S.ws.partials};
  const RhoReq* cycle_rho = fused && nb_cycle_rho > 0 ? &rq : nullptr;
  const int nb_rho = cycle_rho ? nb_cycle_rho : vec_grid(n);
  const int nb_gam = mf ? mf_grid(*mf) : spmv_partials<T>(A, t);
  const double* rho_part = fused ? S.ws.partials : nullptr;
  const double* gam_part = fused ? S.ws.partials : nullptr;
  auto iteration = [&](hipStream_t st) -> int {
    int e;
    const int first = -1;
    if (pkind == 1) {
      if ((e = launch_jacobi_dot<T>(n, k.r, dinv, k.z, S.ws, first, S.flags, st, !fused))) return e;
      if ((e = launch_update_p<T, T>(n, k.z, nullptr, k.p, S.ws, first, S.flags, st, rho_part, nb_rho))) return e;
    } else if (pkind == 2 && amg32) {
      const float* vin = sep32 ? rp32 : reinterpret_cast<const float*>(k.r);
      float* vout = sep32 ? S.zp32 : reinterpret_cast<float*>(k.z);
      if ((e = amg_apply(*amg, amg_vt, vin, vout, S.flags, st, cycle_rho, first_done))) return e;
      if (!cycle_rho &&
          (e = launch_rho<T, float>(n, k.r, vout, scale, S.ws, first, S.flags, st, flexible ? k.q : nullptr, !fused)))
        return e;
      if ((e = launch_update_p<T, float>(n, vout, scale, k.p, S.ws, first, S.flags, st, rho_part, nb_rho))) return e;
    } else if (pkind == 2) {
      // FP64 cycle: in-place scaled input in k.z when S is given
      const T* vin = sep64 ? k.z : k.r;
      T* vout = k.q;  // q is free until the SpMV below
      if ((e = amg_apply(*amg, 0, vin, vout, S.flags, st, cycle_rho, first_done))) return e;
      if (!cycle_rho && (e = launch_rho<T, T>(n, k.r, vout, scale, S.ws, first, S.flags, st, nullptr, !fused)))
        return e;
      if ((e = launch_update_p<T, T>(n, vout, scale, k.p, S.ws, first, S.flags, st, rho_part, nb_rho))) return e;
    } else {
      if ((e = launch_rho<T, T>(n, k.r, k.r, nullptr, S.ws, first, S.flags, st, nullptr, !fused))) return e;
      if ((e = launch_update_p<T, T>(n, k.r, nullptr, k.p, S.ws, first, S.flags, st, rho_part, nb_rho))) return e;
    }
    if (mf) {
      const double* pp = reinterpret_cast<const double*>(k.p);
      double* qq = reinterpret_cast<double*>(k.q);
      if ((e = mf_apply(*mf, pp, nullptr, qq, S.ws.partials, 1, S.flags, 1, st))) return e;
      if (!fused && (e = launch_finalize(S.ws.partials, mf_grid(*mf), 1, kFinGamma, S.ws.sc + kScGamma, S.ws.sc,
                                         nullptr, 0, S.flags, st)))
        return e;
    } else if ((e = spmv_dot<T>(A, k.p, k.q, S.ws, fused ? 2 : 1, t, st))) {
      return e;
    }
    if (sep32)
      return launch_update_xr<T, float>(n, x, k.r, k.p, k.q, rp32, scale, S.ws, S.flags, st, gam_part, nb_gam,
                                        first_done ? &fi32 : nullptr);
    if (sep64)
      return launch_update_xr<T, T>(n, x, k.r, k.p, k.q, k.z, scale, S.ws, S.flags, st, gam_part, nb_gam,
                                    first_done ? &fi64 : nullptr);
    return launch_update_xr<T, T>(n, x, k.r, k.p, k.q, nullptr, nullptr, S.ws, S.flags, st, gam_part, nb_gam);
  };
  std::vector<uint64_t> key;
  if (S.use_graphs) {
    key_push(key, (int)sizeof(T));
    for (int v : {pkind, amg_vt, check_every}) key_push(key, v);
    for (const void* p : {(const void*)dinv, (const void*)scale, (const void*)x, (const void*)k.r, (const void*)k.z,
                          (const void*)k.p, (const void*)k.q, (const void*)rp32, (const void*)S.zp32,
                          (const void*)S.ws.partials, (const void*)S.ws.sc, (const void*)S.flags})
      key_push(key, p);
    key_push(key, n);
    if (mf) {
      key_mf(key, *mf);
    } else {
      key_matrix(key, A);
    }
    key_tune(key, t);
    key_amg(key, pkind == 2 ? amg : nullptr);
  }
  int done_iters = 0;
  // Iterations enqueued after convergence still launch every kernel (each
  // returns at once, but a grid of thousands of workgroups costs several us
  // to retire), so a chunk is sized from the residual reduction rate of the
  // previous one to end near convergence: rounded down, since one more small
  // chunk costs a host round trip, far less than idle iterations.
  int chunk = std::min(check_every, max_iter);
  int it_prev = 0;
  double rr_prev = 0.0;
  for (int it = 0; it < max_iter;) {
    chunk = std::max(1, std::min(chunk, max_iter - it));
    if (S.use_graphs && chunk == check_every) {
      r = graph_for(S, key, [&](hipStream_t st) -> int {
        for (int c = 0; c < chunk; ++c) {
          const int e = iteration(st);
          if (e) return e;
        }
        return kOk;
      });
      if (r) return r;
      WN_HIP(hipGraphLaunch(S.gexec, s));
    } else {
      for (int c = 0; c < chunk; ++c)
        if ((r = iteration(s))) return r;
    }
    it += chunk;
    WN_HIP(hipMemcpyAsync(S.h_flags, S.flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    WN_HIP(hipMemcpyAsync(S.h_sc, S.ws.sc, sizeof(double) * kScCount, hipMemcpyDeviceToHost, s));
    WN_HIP(hipStreamSynchronize(s));
    done_iters = S.h_flags[1];
    if (S.h_flags[0]) break;
    const double rr = S.h_sc[kScRR], tol2 = S.h_sc[kScTol2];
    if (it_prev == 0) rr_prev = S.h_sc[kScRR0];
    chunk = check_every;
    if (rr > 0.0 && rr < rr_prev && tol2 > 0.0 && tol2 < rr) {
      const double need = std::log(tol2 / rr) / (std::log(rr / rr_prev) / (it - it_prev));
      if (need < check_every) chunk = std::max(1, (int)need);
    }
    it_prev = it;
    rr_prev = rr;
  }
  WN_HIP(hipMemcpyAsync(S.h_sc, S.ws.sc, sizeof(double) * kScCount, hipMemcpyDeviceToHost, s));
  WN_HIP(hipStreamSynchronize(s));
  if (iters) *iters = done_iters;
  if (rel) *rel = S.h_sc[kScRR0] > 0.0 ? std::sqrt(S.h_sc[kScRR] / S.h_sc[kScRR0]) : 0.0;
  return kOk;
}
template int pcg_run<double>(Solver&, const SellMatrix&, int, const double*, Amg*, int, const double*, const double*,
                             double*, double, int, int, const SpmvTune&, hipStream_t, int*, double*, const MfOp*);
template int pcg_run<float>(Solver&, const SellMatrix&, int, const float*, Amg*, int, const double*, const float*,
                            float*, double, int, int, const SpmvTune&, hipStream_t, int*, double*, const MfOp*);

// ------------------------------------------------------------------ refinement
// FP64 iterative refinement around an FP32 (equilibrated) inner PCG:
//   r = f - K u (FP64, K = A.d_v64);  rhat = s .* r;  solve Ahat dhat = rhat/|rhat|
//   in FP32 (A.d_v32 = S K S) to relative inner_rtol;  u += |rhat| s .* dhat.
// Stops when ||f - K u|| <= tol ||f||.  Returns total inner iterations.
int ir_run(Solver& S, SellMatrix& A, const double* scale, int pkind, const float* dinv32, Amg* amg, int amg_vt,
           const double* f, double* u, double tol, double inner_rtol, int max_outer, int max_inner, int check_every,
           const SpmvTune& tune, hipStream_t s, int* outer_out, int* inner_out, double* rel_out, int zero_init) {
  const long long n = (long long)A.n_rows * A.br;
  int r;
  if ((r = S.init())) return r;
  if ((r = S.ws.alloc(A.n_slices + 8))) return r;
  if (S.n_ir < n) {
    for (void* p : {(void*)S.r64, (void*)S.rhs32, (void*)S.d32})
      if (p) (void)hipFree(p);
    WN_HIP(hipMalloc(&S.r64, sizeof(double) * n));
    WN_HIP(hipMalloc(&S.rhs32, sizeof(float) * n));
    WN_HIP(hipMalloc(&S.d32, sizeof(float) * n));
    S.n_ir = n;
  }
  if (!A.d_v64 || !A.d_v32) return kNotReady;
  if (zero_init && (r = launch_fill<double>(n, u, 0.0, s))) return r;
  if ((r = launch_dot<double>(n, f, f, S.ws, S.ws.sc + kScF2, s))) return r;
  int total_inner = 0, outer = 0;
  double rel = 0.0;
  SpmvTune t = tune;
  t.flags = nullptr;
  t.guard = 0;
  for (;; ++outer) {
    if ((r = spmv_residual64(A, u, f, S.r64, S.rhs32, scale, S.ws, S.ws.sc + kScTrue2, t, s))) return r;
    WN_HIP(hipMemcpyAsync(S.h_sc, S.ws.sc, sizeof(double) * kScCount, hipMemcpyDeviceToHost, s));
    WN_HIP(hipStreamSynchronize(s));
    const double f2 = S.h_sc[kScF2];
    rel = f2 > 0.0 ? std::sqrt(S.h_sc[kScTrue2] / f2) : 0.0;
    if (rel <= tol || outer >= max_outer || !(rel == rel)) break;
    // normalised FP32 right-hand side
    if ((r = launch_dot<float>(n, S.rhs32, S.rhs32, S.ws, S.ws.sc + kScInnerN, s))) return r;
    if ((r = launch_normalise<float>(n, S.rhs32, S.rhs32, S.ws.sc + kScInnerN, s))) return r;
    int it = 0;
    double irel = 0.0;
    if ((r = pcg_run<float>(S, A, pkind, dinv32, amg, amg_vt, nullptr, S.rhs32, S.d32, inner_rtol, max_inner,
                            check_every, tune, s, &it, &irel)))
      return r;
    total_inner += it;
    if ((r = launch_ir_update<float>(n, u, S.d32, scale, S.ws.sc + kScInnerN, s))) return r;
  }
  if (outer_out) *outer_out = outer;
  if (inner_out) *inner_out = total_inner;
  if (rel_out) *rel_out = rel;
  return kOk;
}

}  // namespace wn

namespace wn {

Solver* solver_new() { return new Solver(); }
void solver_set_graphs(Solver& s, int enable) {
  s.use_graphs = enable ? 1 : 0;
  if (!enable) s.drop_graph();
}
void solver_delete(Solver* s) { delete s; }
Workspace& solver_ws(Solver& s) {
  (void)s.init();
  return s.ws;
}

Amg* amg_new() { return new Amg(); }
void amg_delete(Amg* g) { delete g; }
int amg_resize(Amg& g, int n_levels, int store) {
  g.release();
  g.store = store;
  for (int i = 0; i < n_levels; ++i) g.L.emplace_back(new AmgLevel());
  return kOk;
}

// Level lvl: operator A (level 0: the fine matrix's equilibrated values, given
// by `fine`; other levels: host standard-layout block arrays), point-Jacobi
// inverse diagonal (host), Chebyshev interval [lmax/ratio, lmax] and degree,
// and -- except on the coarsest level -- the prolongator P (br x bc_coarse
// blocks) and restriction R = P^T (bc_coarse x br blocks), host arrays.
int amg_set_level(Amg& g, int lvl, SellMatrix* fine, int br, int n_rows, const int* a_rowptr, const int* a_colidx,
                  const double* a_vals, const double* dinv, double lmax, double ratio, int degree, int bc_coarse,
                  int n_coarse_rows, const int* p_rowptr, const int* p_colidx, const double* p_vals,
                  const int* r_rowptr, const int* r_colidx, const double* r_vals, int sigma) {
  if (lvl < 0 || lvl >= (int)g.L.size() || br <= 0 || n_rows < 0 || degree < 1 || degree > 16 || !(lmax > 0.0) ||
      !(ratio > 1.0))
    return kBadArg;
  AmgLevel& L = *g.L[lvl];
  int r;
  const int to64 = g.store & 1, to32 = (g.store & 6) ? 1 : 0, to16 = (g.store & 4) ? 1 : 0;
  if (fine) {
    if (fine->br != br || fine->n_rows != n_rows || !fine->d_v32) return kBadArg;
    if ((to64 && !fine->d_v64e) || (to16 && !fine->d_v16)) return kNotReady;
    L.A = &fine->equilibrated();
    L.own_A = false;
  } else {
    if (!a_rowptr || !a_vals) return kBadArg;
    L.A = new SellMatrix();
    L.own_A = true;
    if ((r = L.A->build(br, br, n_rows, n_rows, a_rowptr, a_colidx, sigma))) return r;
    double* dv = nullptr;
    const size_t nv = (size_t)L.A->nnzb * br * br;
    WN_HIP(hipMalloc(&dv, sizeof(double) * std::max<size_t>(nv, 1)));
    WN_HIP(hipMemcpy(dv, a_vals, sizeof(double) * nv, hipMemcpyHostToDevice));
    r = sell_from_std(*L.A, dv, to64, to32, nullptr, to16);
    (void)hipDeviceSynchronize();
    (void)hipFree(dv);
    if (r) return r;
  }
  L.n = n_rows * br;
  L.lmax = lmax;
  L.ratio = ratio;
  L.degree = degree;
  if (!L.dinv64) WN_HIP(hipMalloc(&L.dinv64, sizeof(double) * std::max(L.n, 1)));
  if (!L.dinv32) WN_HIP(hipMalloc(&L.dinv32, sizeof(float) * std::max(L.n, 1)));
  std::vector<float> d32(L.n);
  L.unit64 = L.unit32 = true;
  for (int i = 0; i < L.n; ++i) {
    d32[i] = static_cast<float>(dinv[i]);
    L.unit64 = L.unit64 && dinv[i] == 1.0;
    L.unit32 = L.unit32 && d32[i] == 1.0f;
  }
  WN_HIP(hipMemcpy(L.dinv64, dinv, sizeof(double) * L.n, hipMemcpyHostToDevice));
  WN_HIP(hipMemcpy(L.dinv32, d32.data(), sizeof(float) * L.n, hipMemcpyHostToDevice));
  L.has_PR = false;
  if (p_rowptr) {
    if (!p_vals || !r_rowptr || !r_vals || bc_coarse <= 0 || n_coarse_rows < 0) return kBadArg;
    auto upload = [&](SellMatrix& M, int rb, int cb, int rows, int cols, const int* rp, const int* ci,
                      const double* vals) -> int {
      int e = M.build(rb, cb, rows, cols, rp, ci, sigma);
      if (e) return e;
      double* dv = nullptr;
      const size_t nv = (size_t)M.nnzb * rb * cb;
      WN_HIP(hipMalloc(&dv, sizeof(double) * std::max<size_t>(nv, 1)));
      WN_HIP(hipMemcpy(dv, vals, sizeof(double) * nv, hipMemcpyHostToDevice));
      e = sell_from_std(M, dv, to64, to32, nullptr, to16);
      (void)hipDeviceSynchronize();
      (void)hipFree(dv);
      return e;
    };
    if ((r = upload(L.P, br, bc_coarse, n_rows, n_coarse_rows, p_rowptr, p_colidx, p_vals))) return r;
    if ((r = upload(L.R, bc_coarse, br, n_coarse_rows, n_rows, r_rowptr, r_colidx, r_vals))) return r;
    L.has_PR = true;
  }
  return kOk;
}

int amg_set_coarse(Amg& g, int n, const double* inv) {
  // the tile product reads x[n - 1] for its padding lanes: n >= 1
  if (n < 1 || !inv) return kBadArg;
  g.release_coarse();
  const size_t nn = (size_t)n * n;
  WN_HIP(hipMalloc(&g.inv, sizeof(double) * nn));
  WN_HIP(hipMemcpy(g.inv, inv, sizeof(double) * nn, hipMemcpyHostToDevice));
  if (g.store & 6) {
    WN_HIP(hipMalloc(&g.inv32, sizeof(float) * nn));
    const int r = launch_cast_scale<double, float>((long long)nn, g.inv, nullptr, 1.0, g.inv32, nullptr);
    if (r) return r;
    WN_HIP(hipDeviceSynchronize());
  }
  // packed upper-triangle tiles for the single-vector cycle
  g.nt = (n + kTile - 1) / kTile;
  g.ntiles = (long long)g.nt * (g.nt + 1) / 2;
  std::vector<int2> tij((size_t)g.ntiles);
  for (int I = 0; I < g.nt; ++I)
    for (int J = I; J < g.nt; ++J) tij[(size_t)sym_tile_index(I, J, g.nt)] = make_int2(I, J);
  const size_t tsz = (size_t)g.ntiles * kTile * kTile;
  WN_HIP(hipMalloc(&g.tij, sizeof(int2) * tij.size()));
  WN_HIP(hipMemcpy(g.tij, tij.data(), sizeof(int2) * tij.size(), hipMemcpyHostToDevice));
  WN_HIP(hipMalloc(&g.tpart, sizeof(double) * 2 * (size_t)g.ntiles * kTile));
  if (g.store & 1) {
    WN_HIP(hipMalloc(&g.tiles64, sizeof(double) * tsz));
    hipLaunchKernelGGL(sym_tile_pack_kernel<double>, dim3((unsigned)g.ntiles), dim3(256), 0, nullptr, n, g.nt, g.inv,
                       g.tij, g.tiles64);
    WN_HIP(hipGetLastError());
  }
  if (g.store & 6) {
    WN_HIP(hipMalloc(&g.tiles32, sizeof(float) * tsz));
    hipLaunchKernelGGL(sym_tile_pack_kernel<float>, dim3((unsigned)g.ntiles), dim3(256), 0, nullptr, n, g.nt, g.inv,
                       g.tij, g.tiles32);
    WN_HIP(hipGetLastError());
  }
  WN_HIP(hipDeviceSynchronize());
  g.n_coarse = n;
  return kOk;
}

int amg_attach_mf(Amg& g, const MfOp* op, const double* scale, hipStream_t s) {
  if (op && (g.L.empty() || 3LL * op->n_nodes != g.L[0]->n)) return kBadArg;
  return mf_smooth_attach(g.mfs, op, scale, (g.store & 2) ? 1 : 0, s);
}

int amg_set_tune(Amg& g, int prec, int waves, int xcd_remap, int unroll) {
  g.tune.set(prec, waves, xcd_remap, unroll);
  return kOk;
}

int amg_finalize(Amg& g, const SpmvTune& t) {
  g.tune = t;
  for (size_t l = 0; l + 1 < g.L.size(); ++l)
    if (!g.L[l]->A || !g.L[l]->has_PR) return kNotReady;
  if (g.L.empty() || !g.L.back()->A || g.n_coarse != g.L.back()->n) return kNotReady;
  return amg_alloc_vectors(g, g.store);
}

}  // namespace wn
