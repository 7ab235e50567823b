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

// z = dinv .* r, rho = r . z; finalize: beta = rho/rho_prev (0 on the first
// iteration), rho_prev <- rho.  sc layout: see Scalars in wn.hpp.
template <typename T>
__global__ __launch_bounds__(kVecBlock) void jacobi_dot_kernel(long long n, const T* __restrict__ r,
                                                               const T* __restrict__ dinv, T* __restrict__ z,
                                                               double* partials, unsigned int* counter, double* sc,
                                                               int first, const int* skip) {
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
  if (grid_reduce_last<1>(v, partials, counter, sc + kScRho) && threadIdx.x == 0) {
    const double rho = sc[kScRho];
    sc[kScBeta] = first ? 0.0 : (sc[kScRhoPrev] != 0.0 ? rho / sc[kScRhoPrev] : 0.0);
    sc[kScRhoPrev] = rho;
  }
}

// rho = r . z with the same finalisation (the preconditioner wrote z).
template <typename T>
__global__ __launch_bounds__(kVecBlock) void rho_kernel(long long n, const T* __restrict__ r, const T* __restrict__ z,
                                                        double* partials, unsigned int* counter, double* sc,
                                                        int first, const int* skip) {
  if (skip && *skip) return;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += double(r[i]) * double(z[i]);
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  if (grid_reduce_last<1>(v, partials, counter, sc + kScRho) && threadIdx.x == 0) {
    const double rho = sc[kScRho];
    sc[kScBeta] = first ? 0.0 : (sc[kScRhoPrev] != 0.0 ? rho / sc[kScRhoPrev] : 0.0);
    sc[kScRhoPrev] = rho;
  }
}

// p = z + beta p
template <typename T>
__global__ void update_p_kernel(long long n, const T* __restrict__ z, T* __restrict__ p, const double* sc,
                                const int* skip) {
  if (skip && *skip) return;
  const T beta = T(sc[kScBeta]);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// x += alpha p; r -= alpha q; rr = r . r.  Finalise: count the iteration and
// raise the done flag once rr <= sc[kScTol2] (set by the driver to rtol^2 rr0).
template <typename T>
__global__ __launch_bounds__(kVecBlock) void update_xr_kernel(long long n, T* __restrict__ x, T* __restrict__ r,
                                                              const T* __restrict__ p, const T* __restrict__ q,
                                                              double* partials, unsigned int* counter, double* sc,
                                                              int* flags) {
  if (flags && flags[0]) return;
  const T alpha = T(sc[kScAlpha]);
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc += double(ri) * double(ri);
  }
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  if (grid_reduce_last<1>(v, partials, counter, sc + kScRR) && threadIdx.x == 0 && flags) {
    flags[1] += 1;
    if (sc[kScRR] <= sc[kScTol2] || !(sc[kScRR] == sc[kScRR])) flags[0] = 1;
  }
}

// generic dot into out[0]
template <typename T>
__global__ __launch_bounds__(kVecBlock) void dot_kernel(long long n, const T* __restrict__ a, const T* __restrict__ b,
                                                        double* partials, unsigned int* counter, double* out) {
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += double(a[i]) * double(b[i]);
  __shared__ double sm[kVecBlock / 64];
  const double bs = block_sum(acc, sm);
  double v[1] = {bs};
  grid_reduce_last<1>(v, partials, counter, out);
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
                      hipStream_t s) {
  hipLaunchKernelGGL(jacobi_dot_kernel<T>, WN_GRID(n), n, r, dinv, z, w.partials, w.counter, w.sc, first, skip);
  return static_cast<int>(hipGetLastError());
}
template <typename T>
int launch_rho(long long n, const T* r, const T* z, Workspace& w, int first, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(rho_kernel<T>, WN_GRID(n), n, r, z, w.partials, w.counter, w.sc, first, skip);
  return static_cast<int>(hipGetLastError());
}
template <typename T>
int launch_update_p(long long n, const T* z, T* p, Workspace& w, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(update_p_kernel<T>, WN_GRID(n), n, z, p, w.sc, skip);
  return static_cast<int>(hipGetLastError());
}
template <typename T>
int launch_update_xr(long long n, T* x, T* r, const T* p, const T* q, Workspace& w, int* flags, hipStream_t s) {
  hipLaunchKernelGGL(update_xr_kernel<T>, WN_GRID(n), n, x, r, p, q, w.partials, w.counter, w.sc, flags);
  return static_cast<int>(hipGetLastError());
}
template <typename T>
int launch_dot(long long n, const T* a, const T* b, Workspace& w, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_kernel<T>, WN_GRID(n), n, a, b, w.partials, w.counter, out);
  return static_cast<int>(hipGetLastError());
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

#define WN_INST(T)                                                                                   \
  template int launch_jacobi_dot<T>(long long, const T*, const T*, T*, Workspace&, int, const int*, hipStream_t); \
  template int launch_rho<T>(long long, const T*, const T*, Workspace&, int, const int*, hipStream_t);            \
  template int launch_update_p<T>(long long, const T*, T*, Workspace&, const int*, hipStream_t);                  \
  template int launch_update_xr<T>(long long, T*, T*, const T*, const T*, Workspace&, int*, hipStream_t);         \
  template int launch_dot<T>(long long, const T*, const T*, Workspace&, double*, hipStream_t);        \
  template int launch_ir_update<T>(long long, double*, const T*, const double*, const double*, hipStream_t); \
  template int launch_normalise<T>(long long, const T*, T*, const double*, hipStream_t);
WN_INST(float)
WN_INST(double)

}  // namespace wn
