// Sliced-ELL block matrices: host-side layout plan, value conversion and
// assembly kernels, and the SpMV instantiations used by the solvers.
#include <algorithm>
#include <numeric>
#include "wn.hpp"

namespace wn {

int Workspace::alloc(int need) {
  if (need <= max_blocks && partials) return kOk;
  release();
  max_blocks = std::max(need, 4096);
  WN_HIP(hipMalloc(&partials, sizeof(double) * 4 * (size_t)max_blocks));
  WN_HIP(hipMalloc(&sc, sizeof(double) * kScCount));
  WN_HIP(hipMemset(sc, 0, sizeof(double) * kScCount));
  return kOk;
}

void Workspace::release() {
  if (partials) (void)hipFree(partials);
  if (sc) (void)hipFree(sc);
  partials = nullptr;
  sc = nullptr;
  max_blocks = 0;
}

void SellMatrix::release() {
  if (eq_alias) {
    delete eq_alias;  // non-owning view: nothing of its own to free
    eq_alias = nullptr;
  }
  for (void* p : {(void*)d_off, (void*)d_rinfo, (void*)d_col, (void*)d_slot, (void*)d_v64, (void*)d_v64e,
                  (void*)d_v32, (void*)d_v16})
    if (p) (void)hipFree(p);
  d_v16 = nullptr;
  d_off = nullptr;
  d_rinfo = nullptr;
  d_col = nullptr;
  d_slot = nullptr;
  d_v64 = nullptr;
  d_v64e = nullptr;
  d_v32 = nullptr;
}

SellMatrix& SellMatrix::equilibrated() {
  if (!d_v64e) return *this;
  if (!eq_alias) eq_alias = new SellMatrix();
  SellMatrix& a = *eq_alias;
  a.br = br; a.bc = bc; a.n_rows = n_rows; a.n_cols = n_cols; a.n_slices = n_slices; a.sigma = sigma;
  a.lanes = lanes;
  a.nnzb = nnzb; a.slots = slots; a.d_off = d_off; a.d_rinfo = d_rinfo; a.d_col = d_col;
  a.d_slot = d_slot; a.d_v64 = d_v64e; a.d_v64e = nullptr; a.d_v32 = d_v32; a.d_v16 = d_v16;
  return a;
}

// Lanes per block row: enough lanes that short matrices with long rows still
// spread over many wavefronts (about 2048 x 64 lane-rows), and in any case
// as many as keep about six blocks per lane (more row splitting shortens the
// dependent load chain of each lane: the fine C3D10 operator with ~25 blocks
// per row runs 15-20 % faster with 4 lanes per row than with 1), but no more
// than an eighth of the average row length for the first rule.  A 6x6 block
// is four 3x3 blocks' worth of loads, so those rows split further, to about
// two blocks per lane, up to 8 lanes per row (the 6-DOF beam lattice's
// V-cycle: 177 instead of 189 us; beyond 8 lanes the 8-column SpMM of the
// multi-RHS cycle loses more to its wider shuffle reduction than it gains),
// and 3x6 / 6x3 transfer blocks (two 3x3 blocks' worth) to about three per
// lane (the I-beam's prolongation, about six blocks per row: 36.4 -> 34.1 us
// with two lanes per row).
static int auto_lanes(int n_rows, long long nnzb, int bb) {
  const double avg = n_rows > 0 ? double(nnzb) / n_rows : 0.0;
  int L = 1;
  while (L < 64 && (long long)n_rows * L < 64LL * 2048 && L * 8 < avg) L *= 2;
  while (L < 64 && 2 * L * 6 <= avg) L *= 2;
  if (bb >= 36)
    while (L < 8 && 2 * L * 2 <= avg) L *= 2;
  else if (bb >= 18)
    while (L < 8 && 2 * L * 3 <= avg) L *= 2;
  return L;
}

// Host plan: stable sort of rows by length (descending) inside windows of
// `sigma` rows, slices of 64/L consecutive sorted rows, column-major slots.
int SellMatrix::build(int br_, int bc_, int n_rows_, int n_cols_, const int* rowptr, const int* colidx, int sigma_,
                      int lanes_) {
  if (br_ <= 0 || bc_ <= 0 || n_rows_ < 0 || n_cols_ < 0 || !rowptr || (!colidx && rowptr[n_rows_] > 0))
    return kBadArg;
  if (lanes_ > 64 || (lanes_ > 0 && (lanes_ & (lanes_ - 1)))) return kBadArg;
  release();
  br = br_;
  bc = bc_;
  n_rows = n_rows_;
  n_cols = n_cols_;
  sigma = sigma_ <= 0 ? 4096 : sigma_;
  sigma = ((sigma + 63) / 64) * 64;
  nnzb = rowptr[n_rows];
  lanes = lanes_ > 0 ? lanes_ : auto_lanes(n_rows, nnzb, br_ * bc_);
  const int rows_per_slice = 64 / lanes;
  h_rowptr.assign(rowptr, rowptr + n_rows + 1);
  h_colidx.assign(colidx, colidx + nnzb);
  n_slices = (n_rows + rows_per_slice - 1) / rows_per_slice;
  std::vector<int> order(n_rows);
  std::iota(order.begin(), order.end(), 0);
  for (int w0 = 0; w0 < n_rows; w0 += sigma) {
    const int w1 = std::min(n_rows, w0 + sigma);
    std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) {
      return (rowptr[a + 1] - rowptr[a]) > (rowptr[b + 1] - rowptr[b]);
    });
  }
  std::vector<long long> off(n_slices + 1, 0);
  std::vector<int2> rinfo((size_t)n_slices * rows_per_slice, make_int2(-1, 0));
  for (int s = 0; s < n_slices; ++s) {
    int steps = 0;
    for (int r = 0; r < rows_per_slice; ++r) {
      const long long k = (long long)s * rows_per_slice + r;
      if (k >= n_rows) break;
      const int g = order[k];
      const int rl = rowptr[g + 1] - rowptr[g];
      rinfo[(size_t)k] = make_int2(g, rl);
      steps = std::max(steps, (rl + lanes - 1) / lanes);  // sub-lane 0 has the most steps
    }
    off[s + 1] = off[s] + 64LL * steps;
  }
  slots = off[n_slices];
  std::vector<int> col(std::max<long long>(slots, 1), 0);
  std::vector<long long> slot(std::max<long long>(nnzb, 1), 0);
  for (int s = 0; s < n_slices; ++s)
    for (int r = 0; r < rows_per_slice; ++r) {
      const long long k = (long long)s * rows_per_slice + r;
      if (k >= n_rows) break;
      const int g = order[k];
      for (int j = 0; j < rowptr[g + 1] - rowptr[g]; ++j) {
        const long long sl = off[s] + 64LL * (j / lanes) + (j % lanes) * rows_per_slice + r;
        col[sl] = colidx[rowptr[g] + j];
        slot[rowptr[g] + j] = sl;
      }
      // padding slots keep column 0 and value 0; lanes stop at their own
      // step count, so they are never read.
    }
  WN_HIP(hipMalloc(&d_off, sizeof(long long) * off.size()));
  WN_HIP(hipMalloc(&d_rinfo, sizeof(int2) * rinfo.size()));
  WN_HIP(hipMalloc(&d_col, sizeof(int) * col.size()));
  WN_HIP(hipMalloc(&d_slot, sizeof(long long) * slot.size()));
  WN_HIP(hipMemcpy(d_off, off.data(), sizeof(long long) * off.size(), hipMemcpyHostToDevice));
  WN_HIP(hipMemcpy(d_rinfo, rinfo.data(), sizeof(int2) * rinfo.size(), hipMemcpyHostToDevice));
  WN_HIP(hipMemcpy(d_col, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice));
  WN_HIP(hipMemcpy(d_slot, slot.data(), sizeof(long long) * slot.size(), hipMemcpyHostToDevice));
  return kOk;
}

int SellMatrix::ensure_values(int want64, int want32, int want16) {
  const size_t nv = (size_t)std::max<long long>(slots, 1) * br * bc;
  if (want16 && !d_v16) {
    WN_HIP(hipMalloc(&d_v16, sizeof(_Float16) * nv));
    WN_HIP(hipMemset(d_v16, 0, sizeof(_Float16) * nv));
  }
  if (want64 && !d_v64) {
    WN_HIP(hipMalloc(&d_v64, sizeof(double) * nv));
    WN_HIP(hipMemset(d_v64, 0, sizeof(double) * nv));
  }
  if (want32 && !d_v32) {
    WN_HIP(hipMalloc(&d_v32, sizeof(float) * nv));
    WN_HIP(hipMemset(d_v32, 0, sizeof(float) * nv));
  }
  return kOk;
}

__global__ void std_to_sell_kernel(long long nnzb, int bb, const long long* __restrict__ slot,
                                   const double* __restrict__ vs, double* __restrict__ v64, float* __restrict__ v32,
                                   _Float16* __restrict__ v16) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= nnzb * bb) return;
  const long long k = t / bb;
  const int e = (int)(t - k * bb);
  const long long sl = slot[k];
  const double v = vs[t];
  if (v64) v64[val_index<double>(sl, e, bb)] = v;
  if (v32) v32[val_index<float>(sl, e, bb)] = float(v);
  if (v16) v16[val_index<_Float16>(sl, e, bb)] = _Float16(float(v));
}

__global__ void sell_to_std_kernel(long long nnzb, int bb, const long long* __restrict__ slot,
                                   const double* __restrict__ v64, double* __restrict__ vs) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= nnzb * bb) return;
  const long long k = t / bb;
  const int e = (int)(t - k * bb);
  vs[t] = v64[val_index<double>(slot[k], e, bb)];
}

int sell_from_std(SellMatrix& A, const double* vals_std, int to64, int to32, hipStream_t s, int to16) {
  int r = A.ensure_values(to64, to32, to16);
  if (r) return r;
  const long long n = A.nnzb * A.br * A.bc;
  if (n == 0) return kOk;
  hipLaunchKernelGGL(std_to_sell_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, A.nnzb, A.br * A.bc, A.d_slot,
                     vals_std, to64 ? A.d_v64 : nullptr, to32 ? A.d_v32 : nullptr, to16 ? A.d_v16 : nullptr);
  return static_cast<int>(hipGetLastError());
}

int sell_to_std(const SellMatrix& A, double* vals_std, hipStream_t s) {
  if (!A.d_v64) return kNotReady;
  const long long n = A.nnzb * A.br * A.bc;
  if (n == 0) return kOk;
  hipLaunchKernelGGL(sell_to_std_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, A.nnzb, A.br * A.bc, A.d_slot,
                     A.d_v64, vals_std);
  return static_cast<int>(hipGetLastError());
}

// Deterministic assembly of 3x3 blocks.  Block k = (i, j) sums the entries of
// its contributions in list order (element order, then local pair order), each
// code = e*100 + a*10 + b addressing Ke[e][3a+r][3b+c].  Dirichlet semantics of
// solid_solver.py:95-127 per scalar entry: a fixed row or column DOF zeroes the
// entry, and the diagonal entry of a fixed DOF becomes 1.  node_fix[n] holds the
// fixed-DOF bits of node n (bit c = DOF c fixed).
__global__ void assemble_gather_kernel(long long nnzb, const int* __restrict__ rowidx, const int* __restrict__ colidx,
                                       const long long* __restrict__ cptr, const int* __restrict__ codes,
                                       const double* __restrict__ Ke, const unsigned char* __restrict__ node_fix,
                                       const long long* __restrict__ slot, double* __restrict__ v64,
                                       double* __restrict__ vstd) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= nnzb) return;
  double acc[9] = {};
  for (long long q = cptr[k]; q < cptr[k + 1]; ++q) {
    const int code = codes[q];
    const int e = code / 100, ab = code - e * 100, a = ab / 10, b = ab - a * 10;
    const double* kp = Ke + (size_t)e * 900 + (3 * a) * 30 + 3 * b;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[r * 3 + c] += kp[r * 30 + c];
  }
  const int i = rowidx[k], j = colidx[k];
  const unsigned fi = node_fix ? node_fix[i] : 0u, fj = node_fix ? node_fix[j] : 0u;
  const long long sl = slot[k];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = acc[r * 3 + c];
      if (((fi >> r) & 1u) || ((fj >> c) & 1u)) v = (i == j && r == c) ? 1.0 : 0.0;
      v64[val_index<double>(sl, r * 3 + c, 9)] = v;
      if (vstd) vstd[k * 9 + r * 3 + c] = v;
    }
}

int assemble_gather(SellMatrix& A, const int* d_rowidx, const int* d_colidx, const long long* cptr, const int* codes,
                    const double* Ke, const unsigned char* node_fix, double* vals_std, hipStream_t s) {
  if (A.br != 3 || A.bc != 3) return kBadArg;
  int r = A.ensure_values(1, 0);
  if (r) return r;
  if (A.nnzb == 0) return kOk;
  hipLaunchKernelGGL(assemble_gather_kernel, dim3(ceil_div(A.nnzb, 256)), dim3(256), 0, s, A.nnzb, d_rowidx, d_colidx,
                     cptr, codes, Ke, node_fix, A.d_slot, A.d_v64, vals_std);
  return static_cast<int>(hipGetLastError());
}

// Generic version for b x b blocks of npe-node elements (beams: b = 6, npe =
// 2): code = e*npe*npe + a*npe + c addresses Ke[e][b*a + r][b*c + q].
template <int B>
__global__ void assemble_elem_kernel(long long nnzb, int npe, int sym, const int* __restrict__ rowidx,
                                     const int* __restrict__ colidx, const long long* __restrict__ cptr,
                                     const int* __restrict__ codes, const double* __restrict__ Ke,
                                     const unsigned char* __restrict__ node_fix, const long long* __restrict__ slot,
                                     double* __restrict__ v64, double* __restrict__ vstd) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= nnzb) return;
  const int D = npe * B, nn = npe * npe;
  double acc[B * B];
#pragma unroll
  for (int t = 0; t < B * B; ++t) acc[t] = 0.0;
  for (long long q = cptr[k]; q < cptr[k + 1]; ++q) {
    const int code = codes[q];
    const int e = code / nn, ab = code - e * nn, a = ab / npe, b = ab - a * npe;
    const double* kp = Ke + (size_t)e * D * D;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double v = kp[(size_t)(B * a + r) * D + B * b + c];
        if (sym) v = 0.5 * (v + kp[(size_t)(B * b + c) * D + B * a + r]);
        acc[r * B + c] += v;
      }
  }
  const int i = rowidx[k], j = colidx[k];
  const unsigned fi = node_fix ? node_fix[i] : 0u, fj = node_fix ? node_fix[j] : 0u;
  const long long sl = slot[k];
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double v = acc[r * B + c];
      if (((fi >> r) & 1u) || ((fj >> c) & 1u)) v = (i == j && r == c) ? 1.0 : 0.0;
      v64[val_index<double>(sl, r * B + c, B * B)] = v;
      if (vstd) vstd[k * B * B + r * B + c] = v;
    }
}

int assemble_elem(SellMatrix& A, int npe, int sym, const int* d_rowidx, const int* d_colidx, const long long* cptr,
                  const int* codes, const double* Ke, const unsigned char* node_fix, double* vals_std, hipStream_t s) {
  if (A.br != A.bc || npe <= 0 || npe > 27) return kBadArg;
  int r = A.ensure_values(1, 0);
  if (r) return r;
  if (A.nnzb == 0) return kOk;
  const dim3 grid(ceil_div(A.nnzb, 128)), blk(128);
  switch (A.br) {
    case 3:
      hipLaunchKernelGGL(assemble_elem_kernel<3>, grid, blk, 0, s, A.nnzb, npe, sym, d_rowidx, d_colidx, cptr, codes,
                         Ke, node_fix, A.d_slot, A.d_v64, vals_std);
      break;
    case 6:
      hipLaunchKernelGGL(assemble_elem_kernel<6>, grid, blk, 0, s, A.nnzb, npe, sym, d_rowidx, d_colidx, cptr, codes,
                         Ke, node_fix, A.d_slot, A.d_v64, vals_std);
      break;
    default:
      return kBadArg;
  }
  return static_cast<int>(hipGetLastError());
}

__global__ void extract_diag_kernel(int n_rows, int b, const long long* __restrict__ diag_block,
                                    const long long* __restrict__ slot, const double* __restrict__ v64,
                                    double* __restrict__ diag) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)n_rows * b) return;
  const long long i = t / b;
  const int r = (int)(t - i * b);
  const long long k = diag_block[i];
  diag[t] = (k >= 0) ? v64[val_index<double>(slot[k], r * b + r, b * b)] : 0.0;
}

int extract_diag(const SellMatrix& A, const long long* diag_block, double* diag, hipStream_t s) {
  if (A.br != A.bc || !A.d_v64) return kBadArg;
  const long long n = (long long)A.n_rows * A.br;
  if (n == 0) return kOk;
  hipLaunchKernelGGL(extract_diag_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, A.n_rows, A.br, diag_block,
                     A.d_slot, A.d_v64, diag);
  return static_cast<int>(hipGetLastError());
}

// v32 (and v64e) = s_i * v64 * s_j per scalar entry (symmetric diagonal equilibration).
__global__ void equilibrate_kernel(long long nnzb, int b, const int* __restrict__ rowidx,
                                   const int* __restrict__ colidx, const long long* __restrict__ slot,
                                   const double* __restrict__ scale, const double* __restrict__ v64,
                                   float* __restrict__ v32, double* __restrict__ v64e, _Float16* __restrict__ v16) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int bb = b * b;
  if (t >= nnzb * bb) return;
  const long long k = t / bb;
  const int e = (int)(t - k * bb), r = e / b, c = e - r * b;
  const long long sl = slot[k];
  const long long p = val_index<double>(sl, e, bb);
  const double v = scale[(long long)rowidx[k] * b + r] * v64[p] * scale[(long long)colidx[k] * b + c];
  v32[val_index<float>(sl, e, bb)] = float(v);
  if (v64e) v64e[p] = v;
  if (v16) v16[val_index<_Float16>(sl, e, bb)] = _Float16(float(v));
}

int equilibrate(SellMatrix& A, const int* d_rowidx, const int* d_colidx, const double* scale, int want64e,
                hipStream_t s, int want16) {
  if (A.br != A.bc || !A.d_v64) return kBadArg;
  int r = A.ensure_values(0, 1, want16);
  if (r) return r;
  const size_t nv = (size_t)std::max<long long>(A.slots, 1) * A.br * A.bc;
  if (want64e && !A.d_v64e) {
    WN_HIP(hipMalloc(&A.d_v64e, sizeof(double) * nv));
    WN_HIP(hipMemset(A.d_v64e, 0, sizeof(double) * nv));
  }
  const long long n = A.nnzb * A.br * A.bc;
  if (n == 0) return kOk;
  hipLaunchKernelGGL(equilibrate_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, A.nnzb, A.br, d_rowidx, d_colidx,
                     A.d_slot, scale, A.d_v64, A.d_v32, want64e ? A.d_v64e : nullptr, A.d_v16);
  return static_cast<int>(hipGetLastError());
}

// ---------------------------------------------------------------- SpMV dispatch
template <int BR, int BC, typename VT, typename XT, int W, class Op>
static void launch_w(const SellView<VT>& v, const XT* x, const Op& op, int grid, int unroll, int xcd, const int* flags,
                     int guard, hipStream_t s) {
  // keep the unrolled value registers within budget: U*BB values of VT
  constexpr int kMaxU = (BR * BC * (int)sizeof(VT) <= 48) ? 4 : ((BR * BC * (int)sizeof(VT) <= 160) ? 2 : 1);
  const int u = std::min(unroll, kMaxU);
  if (u >= 4 && kMaxU >= 4)
    hipLaunchKernelGGL((sell_spmv_kernel<BR, BC, VT, XT, W, (kMaxU >= 4 ? 4 : 1), Op>), dim3(grid), dim3(64 * W), 0,
                       s, v, x, op, xcd, flags, guard);
  else if (u >= 2 && kMaxU >= 2)
    hipLaunchKernelGGL((sell_spmv_kernel<BR, BC, VT, XT, W, (kMaxU >= 2 ? 2 : 1), Op>), dim3(grid), dim3(64 * W), 0,
                       s, v, x, op, xcd, flags, guard);
  else
    hipLaunchKernelGGL((sell_spmv_kernel<BR, BC, VT, XT, W, 1, Op>), dim3(grid), dim3(64 * W), 0, s, v, x, op, xcd,
                       flags, guard);
}

template <typename VT>
static int sell_grid(const SellMatrix& A, const SpmvTune& t) {
  const int per = (A.n_slices + 7) / 8;
  return t.xcd_for<VT>() ? per * 8 : A.n_slices;
}

template <int BR, int BC, typename VT, typename XT, class Op>
static int launch_sell(const SellMatrix& A, const XT* x, const Op& op, const SpmvTune& t, hipStream_t s) {
  if (A.n_slices == 0) return kOk;
  const SellView<VT> v = A.view<VT>();
  if (!v.val) return kNotReady;
  const int grid = sell_grid<VT>(A, t);
  const int u = t.unroll_for<VT>(), xcd = t.xcd_for<VT>();
  switch (t.waves_for<VT>()) {
    case 2:
      launch_w<BR, BC, VT, XT, 2, Op>(v, x, op, grid, u, xcd, t.flags, t.guard, s);
      break;
    case 4:
      launch_w<BR, BC, VT, XT, 4, Op>(v, x, op, grid, u, xcd, t.flags, t.guard, s);
      break;
    default:
      launch_w<BR, BC, VT, XT, 1, Op>(v, x, op, grid, u, xcd, t.flags, t.guard, s);
  }
  return static_cast<int>(hipGetLastError());
}

#define WN_BLOCK_DISPATCH(A, CALL)                     \
  do {                                                 \
    if ((A).br == 3 && (A).bc == 3) { CALL(3, 3); }    \
    else if ((A).br == 6 && (A).bc == 6) { CALL(6, 6); } \
    else if ((A).br == 3 && (A).bc == 6) { CALL(3, 6); } \
    else if ((A).br == 6 && (A).bc == 3) { CALL(6, 3); } \
    else return kBadArg;                               \
  } while (0)

template <typename VT, typename XT>
int spmv_axpby(const SellMatrix& A, const XT* x, XT* y, double alpha, double beta, const SpmvTune& t, hipStream_t s) {
  using AT = typename AccT<VT>::type;
#define CALL(R, C)                                                                 \
  {                                                                                \
    OpAxpby<R, XT, AT> op{y, AT(alpha), AT(beta)};                                 \
    return launch_sell<R, C, VT, XT>(A, x, op, t, s);                              \
  }
  WN_BLOCK_DISPATCH(A, CALL);
#undef CALL
}
template int spmv_axpby<double, double>(const SellMatrix&, const double*, double*, double, double, const SpmvTune&,
                                        hipStream_t);
template int spmv_axpby<float, float>(const SellMatrix&, const float*, float*, double, double, const SpmvTune&,
                                      hipStream_t);

template <typename VT>
int spmv_dot(const SellMatrix& A, const VT* p, VT* y, Workspace& w, int with_alpha, const SpmvTune& t,
             hipStream_t s) {
  using AT = typename AccT<VT>::type;
  if (A.br != A.bc) return kBadArg;
  int r = w.alloc(A.n_slices + 8);
  if (r) return r;
#define CALL(R, C)                                       \
  {                                                      \
    OpSpmvDot<R, VT, AT> op{y, p, w.partials};           \
    r = launch_sell<R, C, VT, VT>(A, p, op, t, s);       \
  }
  if (A.br == 3) CALL(3, 3) else if (A.br == 6) CALL(6, 6) else return kBadArg;
#undef CALL
  if (r || A.n_slices == 0 || with_alpha == 2) return r;
  return launch_finalize(w.partials, sell_grid<VT>(A, t), 1, with_alpha ? kFinGamma : kFinStore, w.sc + kScGamma, w.sc,
                         nullptr, 0, t.guard ? t.flags : nullptr, s);
}
template <typename VT>
int spmv_partials(const SellMatrix& A, const SpmvTune& t) {
  return sell_grid<VT>(A, t);
}
template int spmv_partials<double>(const SellMatrix&, const SpmvTune&);
template int spmv_partials<float>(const SellMatrix&, const SpmvTune&);
template int spmv_dot<double>(const SellMatrix&, const double*, double*, Workspace&, int, const SpmvTune&,
                              hipStream_t);
template int spmv_dot<float>(const SellMatrix&, const float*, float*, Workspace&, int, const SpmvTune&, hipStream_t);

int spmv_residual64(const SellMatrix& A, const double* x, const double* b, double* r, float* rs, const double* scale,
                    Workspace& w, double* out, const SpmvTune& t, hipStream_t s) {
  if (A.br != A.bc) return kBadArg;
  int e = w.alloc(A.n_slices + 8);
  if (e) return e;
#define CALL(R, C)                                                  \
  {                                                                 \
    OpResidual<R, double> op{b, r, rs, scale, w.partials};          \
    e = launch_sell<R, C, double, double>(A, x, op, t, s);          \
  }
  if (A.br == 3) CALL(3, 3) else if (A.br == 6) CALL(6, 6) else return kBadArg;
#undef CALL
  if (e || A.n_slices == 0) return e;
  return launch_finalize(w.partials, sell_grid<double>(A, t), 2, kFinStore, out, w.sc, nullptr, 0,
                         t.guard ? t.flags : nullptr, s);
}

template <typename VT, typename T>
int spmv_cheb(const SellMatrix& A, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, double c1, double c2,
              const SpmvTune& t, hipStream_t s) {
  if (A.br != A.bc) return kBadArg;
#define CALL(R, C)                                                  \
  {                                                                 \
    OpCheb<R, T> op{b, dinv, x_in, x_out, d, T(c1), T(c2)};         \
    return launch_sell<R, C, VT, T>(A, x_in, op, t, s);             \
  }
  if (A.br == 3) CALL(3, 3) else if (A.br == 6) CALL(6, 6) else return kBadArg;
#undef CALL
}

template <typename VT, typename T>
int spmv_defect(const SellMatrix& A, const T* x, const T* b, T* res, const SpmvTune& t, hipStream_t s) {
  if (A.br != A.bc) return kBadArg;
#define CALL(R, C)                                     \
  {                                                    \
    OpDefect<R, T> op{b, res};                         \
    return launch_sell<R, C, VT, T>(A, x, op, t, s);   \
  }
  if (A.br == 3) CALL(3, 3) else if (A.br == 6) CALL(6, 6) else return kBadArg;
#undef CALL
}

template <typename VT, typename T>
int spmv_addto(const SellMatrix& A, const T* e, T* x, const SpmvTune& t, hipStream_t s) {
#define CALL(R, C)                                   \
  {                                                  \
    OpAddTo<R, T> op{x};                             \
    return launch_sell<R, C, VT, T>(A, e, op, t, s); \
  }
  WN_BLOCK_DISPATCH(A, CALL);
#undef CALL
}

template <typename VT, typename T>
int spmv_restrict(const SellMatrix& A, const T* x, T* y, const SpmvTune& t, hipStream_t s) {
  using AT = typename AccT<VT>::type;
#define CALL(R, C)                                   \
  {                                                  \
    OpAxpby<R, T, AT> op{y, AT(1), AT(0)};           \
    return launch_sell<R, C, VT, T>(A, x, op, t, s); \
  }
  WN_BLOCK_DISPATCH(A, CALL);
#undef CALL
}

template <typename VT, typename T>
int spmv_restrict_first(const SellMatrix& A, const T* x, T* y, const T* dinv, T* x1, T* d1, double c2,
                        const SpmvTune& t, hipStream_t s) {
  using AT = typename AccT<VT>::type;
#define CALL(R, C)                                         \
  {                                                        \
    OpRestrictFirst<R, T, AT> op{y, dinv, x1, d1, T(c2)};  \
    return launch_sell<R, C, VT, T>(A, x, op, t, s);       \
  }
  WN_BLOCK_DISPATCH(A, CALL);
#undef CALL
}

#define WN_VINST(VT, T)                                                                                          \
  template int spmv_cheb<VT, T>(const SellMatrix&, const T*, T*, const T*, const T*, T*, double, double,       \
                                const SpmvTune&, hipStream_t);                                                 \
  template int spmv_defect<VT, T>(const SellMatrix&, const T*, const T*, T*, const SpmvTune&, hipStream_t);   \
  template int spmv_addto<VT, T>(const SellMatrix&, const T*, T*, const SpmvTune&, hipStream_t);              \
  template int spmv_restrict<VT, T>(const SellMatrix&, const T*, T*, const SpmvTune&, hipStream_t);           \
  template int spmv_restrict_first<VT, T>(const SellMatrix&, const T*, T*, const T*, T*, T*, double,            \
                                          const SpmvTune&, hipStream_t);
WN_VINST(double, double)
WN_VINST(float, float)
WN_VINST(_Float16, float)

}  // namespace wn
