// Vector kernels of the Krylov loops.  Every dot product is reduced in a fixed
// order (per-block partials, then the last block sums them in index order),
// so results are bit-reproducible run to run.  Scalars (rho, alpha, beta, ...)
// live in device memory and are produced/consumed by the kernels themselves,
// so an iteration needs no host round trip.
#include "wn.hpp"

namespace wn {

constexpr int kVecBlock = 256;

int vec_grid(long long n) {
  // enough blocks to fill 256 CUs several times, capped so the partial
  // arrays stay small; grid-stride loops cover the rest
  long long g = (n + kVecBlock - 1) / kVecBlock;
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

template <typename T>
__global__ void fill_kernel(long long n, T* x, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = v;
}

template <typename A, typename B>
__global__ void cast_scale_kernel(long long n, const A* __restrict__ x, const double* __restrict__ s, double f,
                                  B* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v = double(x[i]) * f;
    if (s) v *= s[i];
    y[i] = B(v);
  }
}

// z = dinv .* r and the partials of rho = r . z (finalised with kFinRho).
template <typename T>
__global__ __launch_bounds__(kVecBlock) void jacobi_dot_kernel(long long n, const T* __restrict__ r,
                                                               const T* __restrict__ dinv, T* __restrict__ z,
                                                               double* partials, const int* skip) {
  if (skip && *skip) return;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const T zi = dinv[i] * r[i];
    z[i] = zi;
    acc += double(r[i]) * double(zi);
  }
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  store_block_partials<1>(v, partials);
}

// partials of rho = r . (S z) (finalised with kFinRho); z may be in a lower
// precision than r (preconditioner output) and S (per-DOF scale) may be null.
// With q (flexible PCG) also the partials of (S z) . q, q = A p of the
// previous iteration (finalised with kFinRhoFlex).
template <typename T, typename ZT>
__global__ __launch_bounds__(kVecBlock) void rho_kernel(long long n, const T* __restrict__ r, const ZT* __restrict__ z,
                                                        const double* __restrict__ scale, const T* __restrict__ q,
                                                        double* partials, const int* skip) {
  if (skip && *skip) return;
  double acc = 0.0, acq = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zs = double(z[i]) * (scale ? scale[i] : 1.0);
    acc += double(r[i]) * zs;
    if (q) acq += zs * double(q[i]);
  }
  __shared__ double sm[kVecBlock / 64];
  double v[2];
  v[0] = block_sum(acc, sm);
  if (q) {
    v[1] = block_sum(acq, sm);
    store_block_partials<2>(v, partials);
  } else {
    double v1[1] = {v[0]};
    store_block_partials<1>(v1, partials);
  }
}

// Sum of nb partials, the same bits in every block of a launch (same
// threads, same order): a consumer kernel finalises a reduction itself
// instead of a one-block launch before it.  A thread adds partials
// t, t + B, t + 2B, ... in index order; the loads of a round of 8 are issued
// together, so the prologue costs one memory latency per 8 B partials
// instead of one per B.
__device__ __forceinline__ double block_partial_sum(const double* __restrict__ part, int nb, double* sm) {
  constexpr int B = kVecBlock, U = 8;
  double a = 0.0;
  for (int base = 0; base < nb; base += U * B) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * B + (int)threadIdx.x;
      v[u] = i < nb ? part[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) a += v[u];
  }
  return block_sum(a, sm);
}

// p = S z + beta p  (p = S z on the first iteration).  first < 0: the first
// iteration is the one where the loop's iteration counter skip[1] is still 0,
// so a captured iteration can be replayed unchanged.  With rho_part every
// block sums the nb_rho partials of rho = r . z itself and takes
// beta = rho / rho_prev; block 0 stores rho and beta (rho_prev is left for
// the update of x and r, after every block here has read it).
template <typename T, typename ZT>
__global__ __launch_bounds__(kVecBlock) void update_p_kernel(long long n, const ZT* __restrict__ z,
                                                             const double* __restrict__ scale, T* __restrict__ p,
                                                             double* sc, int first, const int* skip,
                                                             const double* __restrict__ rho_part, int nb_rho) {
  if (skip && *skip) return;
  const bool f = first >= 0 ? first != 0 : skip[1] == 0;
  double beta_d = sc[kScBeta];
  if (rho_part) {
    __shared__ double sm[kVecBlock / 64];
    const double rho = block_partial_sum(rho_part, nb_rho, sm);
    const double prev = sc[kScRhoPrev];
    beta_d = f ? 0.0 : (prev != 0.0 ? rho / prev : 0.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[kScRho] = rho;
      sc[kScBeta] = beta_d;
    }
  }
  const T beta = T(beta_d);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const T zi = scale ? T(scale[i] * double(z[i])) : T(z[i]);
    p[i] = f ? zi : zi + beta * p[i];
  }
}

// x += alpha p; r -= alpha q; partials of rr = r . r (finalised with kFinRR);
// optionally rp = S r in the preconditioner's precision.  With gam_part
// every block sums the nb_gam partials of gamma = p . A p itself and takes
// alpha = rho / gamma; block 0 stores gamma, alpha and rho_prev = rho.
template <typename T, typename RT>
__global__ __launch_bounds__(kVecBlock) void update_xr_kernel(long long n, T* __restrict__ x, T* __restrict__ r,
                                                              const T* __restrict__ p, const T* __restrict__ q,
                                                              RT* __restrict__ rp, const double* __restrict__ scale,
                                                              double* partials, double* sc, const int* flags,
                                                              const double* __restrict__ gam_part, int nb_gam,
                                                              const RT* __restrict__ fdinv, RT* __restrict__ fx1,
                                                              RT* __restrict__ fd1, RT fc2) {
  if (flags && flags[0]) return;
  double alpha_d = sc[kScAlpha];
  if (gam_part) {
    __shared__ double sg[kVecBlock / 64];
    const double gam = block_partial_sum(gam_part, nb_gam, sg);
    const double rho = sc[kScRho];
    alpha_d = gam != 0.0 ? rho / gam : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc[kScGamma] = gam;
      sc[kScAlpha] = alpha_d;
      sc[kScRhoPrev] = rho;
    }
  }
  const T alpha = T(alpha_d);
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    if (rp) {
      const RT b = RT(scale ? scale[i] * double(ri) : double(ri));
      rp[i] = b;
      if (fx1) {  // the next cycle's first iterate, as the first-step kernel forms it
        const RT v = fc2 * (fdinv ? fdinv[i] : RT(1)) * b;
        fd1[i] = v;
        fx1[i] = v;
      }
    }
    acc += double(ri) * double(ri);
  }
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  store_block_partials<1>(v, partials);
}

// generic dot into out[0]
template <typename T>
__global__ __launch_bounds__(kVecBlock) void dot_kernel(long long n, const T* __restrict__ a, const T* __restrict__ b,
                                                        double* partials) {
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += double(a[i]) * double(b[i]);
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  store_block_partials<1>(v, partials);
}

// u += f * s .* d   (FP64 refinement update; d in working precision)
template <typename T>
__global__ void ir_update_kernel(long long n, double* __restrict__ u, const T* __restrict__ d,
                                 const double* __restrict__ s, const double* sc_norm) {
  const double f = sqrt(*sc_norm);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    u[i] += f * (s ? s[i] : 1.0) * double(d[i]);
}

// y = x / sqrt(*nrm2) (normalised right-hand side of an inner solve)
template <typename T>
__global__ void normalise_kernel(long long n, const T* __restrict__ x, T* __restrict__ y, const double* nrm2) {
  const double g = *nrm2;
  const double f = g > 0.0 ? 1.0 / sqrt(g) : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = T(double(x[i]) * f);
}

// One-block finalisation of a two-launch reduction (see common.hpp).
__global__ __launch_bounds__(1024) void dot_finalize_kernel(const double* __restrict__ partials, int nb, int K,
                                                            int mode, double* out, double* sc, int* flags, int first,
                                                            const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[2][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int k = 0; k < K; ++k) {
    // four independent accumulators keep several loads in flight per thread
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
  if (threadIdx.x != 0) return;
  double v[2] = {0.0, 0.0};
  for (int k = 0; k < K; ++k)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v[k] += red[k][w];
  switch (mode) {
    case kFinRho: {
      sc[kScRho] = v[0];
      const double prev = sc[kScRhoPrev];
      const bool f = first >= 0 ? first != 0 : skip[1] == 0;
      sc[kScBeta] = f ? 0.0 : (prev != 0.0 ? v[0] / prev : 0.0);
      sc[kScRhoPrev] = v[0];
      break;
    }
    case kFinRhoFlex: {
      // beta = z_new . (r_new - r_old) / rho_old = -alpha_old (z_new . q_old) / rho_old
      sc[kScRho] = v[0];
      const double prev = sc[kScRhoPrev];
      const bool f = first >= 0 ? first != 0 : skip[1] == 0;
      sc[kScBeta] = f ? 0.0 : (prev != 0.0 ? -sc[kScAlpha] * v[1] / prev : 0.0);
      sc[kScRhoPrev] = v[0];
      break;
    }
    case kFinGamma:
      sc[kScGamma] = v[0];
      sc[kScAlpha] = v[0] != 0.0 ? sc[kScRho] / v[0] : 0.0;
      break;
    case kFinRR:
      sc[kScRR] = v[0];
      if (flags) {
        flags[1] += 1;
        if (v[0] <= sc[kScTol2] || !(v[0] == v[0])) flags[0] = 1;
      }
      break;
    default:
      for (int k = 0; k < K; ++k) out[k] = v[k];
  }
}

int launch_finalize(const double* partials, int nb, int K, int mode, double* out, double* sc, int* flags, int first,
                    const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(dot_finalize_kernel, dim3(1), dim3(nb > 4096 ? 1024 : 256), 0, s, partials, nb, K, mode, out, sc,
                     flags, first, skip);
  return static_cast<int>(hipGetLastError());
}

// ---------------------------------------------------------------- launchers
#define WN_GRID(n) dim3(vec_grid(n)), dim3(kVecBlock), 0, s

template <typename T>
int launch_fill(long long n, T* x, T v, hipStream_t s) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL(fill_kernel<T>, WN_GRID(n), n, x, v);
  return static_cast<int>(hipGetLastError());
}
template int launch_fill<float>(long long, float*, float, hipStream_t);
template int launch_fill<double>(long long, double*, double, hipStream_t);
template int launch_fill<int>(long long, int*, int, hipStream_t);

template <typename A, typename B>
int launch_cast_scale(long long n, const A* x, const double* sc, double f, B* y, hipStream_t s) {
  if (n <= 0) return kOk;
  hipLaunchKernelGGL((cast_scale_kernel<A, B>), WN_GRID(n), n, x, sc, f, y);
  return static_cast<int>(hipGetLastError());
}
template int launch_cast_scale<double, float>(long long, const double*, const double*, double, float*, hipStream_t);
template int launch_cast_scale<float, double>(long long, const float*, const double*, double, double*, hipStream_t);
template int launch_cast_scale<double, double>(long long, const double*, const double*, double, double*,
                                               hipStream_t);
template int launch_cast_scale<float, float>(long long, const float*, const double*, double, float*, hipStream_t);

template <typename T>
int launch_jacobi_dot(long long n, const T* r, const T* dinv, T* z, Workspace& w, int first, const int* skip,
                      hipStream_t s, bool fin) {
  const int g = vec_grid(n);
  hipLaunchKernelGGL(jacobi_dot_kernel<T>, dim3(g), dim3(kVecBlock), 0, s, n, r, dinv, z, w.partials, skip);
  int e = static_cast<int>(hipGetLastError());
  return e || !fin ? e : launch_finalize(w.partials, g, 1, kFinRho, nullptr, w.sc, nullptr, first, skip, s);
}
template <typename T, typename ZT>
int launch_rho(long long n, const T* r, const ZT* z, const double* scale, Workspace& w, int first, const int* skip,
               hipStream_t s, const T* q, bool fin) {
  const int g = vec_grid(n);
  hipLaunchKernelGGL((rho_kernel<T, ZT>), dim3(g), dim3(kVecBlock), 0, s, n, r, z, scale, q, w.partials, skip);
  int e = static_cast<int>(hipGetLastError());
  if (e || (!fin && !q)) return e;
  return launch_finalize(w.partials, g, q ? 2 : 1, q ? kFinRhoFlex : kFinRho, nullptr, w.sc, nullptr, first, skip, s);
}
template <typename T, typename ZT>
int launch_update_p(long long n, const ZT* z, const double* scale, T* p, Workspace& w, int first, const int* skip,
                    hipStream_t s, const double* rho_part, int nb_rho) {
  hipLaunchKernelGGL((update_p_kernel<T, ZT>), WN_GRID(n), n, z, scale, p, w.sc, first, skip, rho_part, nb_rho);
  return static_cast<int>(hipGetLastError());
}
template <typename T, typename RT>
int launch_update_xr(long long n, T* x, T* r, const T* p, const T* q, RT* rp, const double* scale, Workspace& w,
                     int* flags, hipStream_t s, const double* gam_part, int nb_gam, const FirstIter<RT>* fi) {
  const int g = vec_grid(n);
  // the rr partials go past the region the gamma partials may occupy
  double* rr_part = gam_part ? w.partials + 2 * (size_t)w.max_blocks : w.partials;
  const bool f = fi && rp && fi->x1;
  hipLaunchKernelGGL((update_xr_kernel<T, RT>), dim3(g), dim3(kVecBlock), 0, s, n, x, r, p, q, rp, scale, rr_part,
                     w.sc, flags, gam_part, nb_gam, f && !fi->unit ? fi->dinv : nullptr, f ? fi->x1 : nullptr,
                     f ? fi->d1 : nullptr, f ? fi->c2 : RT(0));
  int e = static_cast<int>(hipGetLastError());
  return e ? e : launch_finalize(rr_part, g, 1, kFinRR, nullptr, w.sc, flags, 0, flags, s);
}
template <typename T>
int launch_dot(long long n, const T* a, const T* b, Workspace& w, double* out, hipStream_t s) {
  const int g = vec_grid(n);
  hipLaunchKernelGGL(dot_kernel<T>, dim3(g), dim3(kVecBlock), 0, s, n, a, b, w.partials);
  int e = static_cast<int>(hipGetLastError());
  return e ? e : launch_finalize(w.partials, g, 1, kFinStore, out, w.sc, nullptr, 0, nullptr, s);
}
template <typename T>
int launch_ir_update(long long n, double* u, const T* d, const double* sc_vec, const double* nrm2, hipStream_t s) {
  hipLaunchKernelGGL(ir_update_kernel<T>, WN_GRID(n), n, u, d, sc_vec, nrm2);
  return static_cast<int>(hipGetLastError());
}
template <typename T>
int launch_normalise(long long n, const T* x, T* y, const double* nrm2, hipStream_t s) {
  hipLaunchKernelGGL(normalise_kernel<T>, WN_GRID(n), n, x, y, nrm2);
  return static_cast<int>(hipGetLastError());
}

// HBM streaming-read probe: every thread sums the floats of eight 16-byte
// non-temporal loads per round (grid-stride over the buffer) and stores its
// sum.  It measures the read rate a plain stream reaches on this device, the
// practical ceiling for the bandwidth-bound kernels; out[t] summed over all
// threads is the sum of the buffer.
constexpr int kStreamBlock = 512, kStreamU = 8;
__global__ __launch_bounds__(kStreamBlock) void stream_read_kernel(const float* __restrict__ a, long long n4,
                                                                   float* __restrict__ out) {
  using F4 = float __attribute__((ext_vector_type(4)));
  const F4* __restrict__ v = reinterpret_cast<const F4*>(a);
  const long long stride = (long long)gridDim.x * kStreamBlock;
  long long i = blockIdx.x * (long long)kStreamBlock + threadIdx.x;
  float s = 0.f;
  for (; i + (kStreamU - 1) * stride < n4; i += kStreamU * stride) {
    F4 w[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) w[u] = __builtin_nontemporal_load(v + i + u * stride);
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) s += (w[u].x + w[u].y) + (w[u].z + w[u].w);
  }
  for (; i < n4; i += stride) {
    const F4 w = __builtin_nontemporal_load(v + i);
    s += (w.x + w.y) + (w.z + w.w);
  }
  out[blockIdx.x * (long long)kStreamBlock + threadIdx.x] = s;
}

int launch_stream_read(const float* a, long long bytes, float* out, int grid, hipStream_t s) {
  if (!a || !out || bytes < 0 || bytes % 16 || grid <= 0 || (reinterpret_cast<uintptr_t>(a) & 15)) return kBadArg;
  hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(kStreamBlock), 0, s, a, bytes / 16, out);
  return static_cast<int>(hipGetLastError());
}

#define WN_INST(T)                                                                                   \
  template int launch_jacobi_dot<T>(long long, const T*, const T*, T*, Workspace&, int, const int*, hipStream_t, \
                                    bool);                                                           \
  template int launch_dot<T>(long long, const T*, const T*, Workspace&, double*, hipStream_t);        \
  template int launch_ir_update<T>(long long, double*, const T*, const double*, const double*, hipStream_t); \
  template int launch_normalise<T>(long long, const T*, T*, const double*, hipStream_t);
WN_INST(float)
WN_INST(double)

#define WN_INST2(T, U)                                                                                           \
  template int launch_rho<T, U>(long long, const T*, const U*, const double*, Workspace&, int, const int*,      \
                                hipStream_t, const T*, bool);                                                  \
  template int launch_update_p<T, U>(long long, const U*, const double*, T*, Workspace&, int, const int*,      \
                                     hipStream_t, const double*, int);                                         \
  template int launch_update_xr<T, U>(long long, T*, T*, const T*, const T*, U*, const double*, Workspace&,    \
                                      int*, hipStream_t, const double*, int, const FirstIter<U>*);
WN_INST2(double, double)
WN_INST2(double, float)
WN_INST2(float, float)

}  // namespace wn
