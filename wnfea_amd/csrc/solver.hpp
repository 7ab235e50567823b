// Host-side state of the AMG hierarchy and of the Krylov solvers (shared by
// solver.hip and multi.hip).
#pragma once
#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>
#include "wn.hpp"

namespace wn {

struct AmgLevel {
  SellMatrix* A = nullptr;  // not owned for level 0
  bool own_A = false;
  SellMatrix P, R;
  bool has_PR = false;
  int n = 0;  // scalar rows
  double* dinv64 = nullptr;
  float* dinv32 = nullptr;
  // the inverse diagonal is exactly 1 in that precision (an equilibrated
  // level): matrix-free smoothing steps skip its load
  bool unit64 = false, unit32 = false;
  double lmax = 2.0, ratio = 30.0;
  int degree = 2;
  // level vectors (b, x, tmp, d, res) per arithmetic precision
  double* v64[5] = {};
  float* v32[5] = {};
};

struct Amg {
  std::vector<std::unique_ptr<AmgLevel>> L;
  int store = 3;  // bit0 FP64, bit1 FP32, bit2 FP16 values
  int n_coarse = 0;
  double* inv = nullptr;   // n_coarse^2, row-major
  float* inv32 = nullptr;  // FP32 copy for the FP32/FP16 cycles (when FP32 values are stored)
  // upper-triangle 64 x 64 tiles of the inverse (symmetric product), their
  // (I, J) table and the FP64 partial products
  int nt = 0;
  long long ntiles = 0;
  double* tiles64 = nullptr;
  float* tiles32 = nullptr;
  int2* tij = nullptr;
  double* tpart = nullptr;
  int coarse_fp64 = 0;  // FP32 cycles apply the FP64 inverse tiles instead of the FP32 copy
  SpmvTune tune;
  // optional: the finest level's smoothing and defect products matrix-free
  MfSmooth mfs;
  ~Amg() { release(); }
  void release() {
    mf_smooth_release(mfs);
    for (auto& lp : L) {
      AmgLevel& l = *lp;
      if (l.own_A && l.A) {
        l.A->release();
        delete l.A;
      }
      l.P.release();
      l.R.release();
      for (auto* p : l.v64) if (p) (void)hipFree(p);
      for (auto* p : l.v32) if (p) (void)hipFree(p);
      if (l.dinv64) (void)hipFree(l.dinv64);
      if (l.dinv32) (void)hipFree(l.dinv32);
    }
    L.clear();
    release_coarse();
  }
  void release_coarse() {
    for (void* p : {(void*)inv, (void*)inv32, (void*)tiles64, (void*)tiles32, (void*)tij, (void*)tpart})
      if (p) (void)hipFree(p);
    inv = nullptr;
    inv32 = nullptr;
    tiles64 = nullptr;
    tiles32 = nullptr;
    tij = nullptr;
    tpart = nullptr;
    nt = 0;
    ntiles = 0;
    n_coarse = 0;
  }
};

template <typename T> T** level_vecs(AmgLevel& l);
template <> inline double** level_vecs<double>(AmgLevel& l) { return l.v64; }
template <> inline float** level_vecs<float>(AmgLevel& l) { return l.v32; }
template <typename T> T* level_dinv(AmgLevel& l);
template <> inline double* level_dinv<double>(AmgLevel& l) { return l.dinv64; }
template <> inline float* level_dinv<float>(AmgLevel& l) { return l.dinv32; }
template <typename T> inline bool level_unit(const AmgLevel& l);
template <> inline bool level_unit<double>(const AmgLevel& l) { return l.unit64; }
template <> inline bool level_unit<float>(const AmgLevel& l) { return l.unit32; }

struct ChebCoef {
  double c1[16], c2[16];
};

inline ChebCoef cheb_coefs(double lmax, double ratio, int deg) {
  ChebCoef c{};
  const double lmin = lmax / ratio;
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sig = th / de;
  double rho = 1.0 / sig;
  c.c1[0] = 0.0;
  c.c2[0] = 1.0 / th;
  for (int k = 1; k < deg && k < 16; ++k) {
    const double rn = 1.0 / (2.0 * sig - rho);
    c.c1[k] = rn * rho;
    c.c2[k] = 2.0 * rn / de;
    rho = rn;
  }
  return c;
}

// VT: values of this level; VTC: values of the levels below it.
template <typename T>
struct KVec {
  T *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  long long n = 0;
  int ensure(long long n_) {
    if (n_ <= n && r) return kOk;
    release();
    n = n_;
    const size_t bytes = sizeof(T) * std::max<long long>(n, 1);
    WN_HIP(hipMalloc(&r, bytes));
    WN_HIP(hipMalloc(&z, bytes));
    WN_HIP(hipMalloc(&p, bytes));
    WN_HIP(hipMalloc(&q, bytes));
    return kOk;
  }
  void release() {
    for (T* v : {r, z, p, q})
      if (v) (void)hipFree(v);
    r = z = p = q = nullptr;
    n = 0;
  }
};

struct Solver {
  Workspace ws;
  int* flags = nullptr;  // [0] done, [1] iterations
  int* h_flags = nullptr;
  double* h_sc = nullptr;
  KVec<double> k64;
  KVec<float> k32;
  // preconditioner input/output when it runs in FP32 under an FP64 Krylov loop
  float* rp32 = nullptr;
  float* zp32 = nullptr;
  long long n_p = 0;
  double* r64 = nullptr;  // FP64 residual of the refinement loop
  float* rhs32 = nullptr;
  float* d32 = nullptr;
  long long n_ir = 0;
  // Replayable chunk of PCG iterations (HIP graph) and the key of everything
  // its kernel arguments were taken from: a solve with the same key replays
  // it, any other solve captures a new one.
  int use_graphs = 0;
  // multi-right-hand-side workspace (multi.hip); freed through multi_free
  void* multi = nullptr;
  void (*multi_free)(void*) = nullptr;
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gkey;
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gkey.clear();
  }
  int init() {
    if (flags) return kOk;
    WN_HIP(hipMalloc(&flags, sizeof(int) * 4));
    WN_HIP(hipMemset(flags, 0, sizeof(int) * 4));
    WN_HIP(hipHostMalloc(&h_flags, sizeof(int) * 4, 0));
    WN_HIP(hipHostMalloc(&h_sc, sizeof(double) * kScCount, 0));
    return ws.alloc(4096);
  }
  int ensure_p(long long n) {
    if (n <= n_p && rp32) return kOk;
    if (rp32) (void)hipFree(rp32);
    if (zp32) (void)hipFree(zp32);
    WN_HIP(hipMalloc(&rp32, sizeof(float) * std::max<long long>(n, 1)));
    WN_HIP(hipMalloc(&zp32, sizeof(float) * std::max<long long>(n, 1)));
    n_p = n;
    return kOk;
  }
  ~Solver() {
    if (multi && multi_free) multi_free(multi);
    drop_graph();
    if (cap) (void)hipStreamDestroy(cap);
    ws.release();
    k64.release();
    k32.release();
    if (flags) (void)hipFree(flags);
    if (h_flags) (void)hipHostFree(h_flags);
    if (h_sc) (void)hipHostFree(h_sc);
    for (void* p : {(void*)r64, (void*)rhs32, (void*)d32, (void*)rp32, (void*)zp32})
      if (p) (void)hipFree(p);
  }
};

int amg_alloc_vectors(Amg& g, int prec_mask);
// One V-cycle z = M r of kind vt; with rq also the partials of rho = r . (S z)
// (amg_rho_blocks(g, vt) of them; kBadMode when that is 0).
// first_done: the finest level's first Chebyshev iterate is in place already
// (see amg_first_iter).
int amg_apply(Amg& g, int vt, const void* r, void* z, const int* skip, hipStream_t s, const RhoReq* rq = nullptr,
              bool first_done = false);
int amg_rho_blocks(const Amg& g, int vt);
// Where the finest level's first Chebyshev iterate of a cycle of kind vt
// (0: T double, 1: T float) with output z goes, so that the kernel writing
// the cycle's input can form it (false: no such cycle).
template <typename T>
bool amg_first_iter(Amg& g, int vt, T* z, FirstIter<T>& f);
// That first iterate from the input b, as the cycle's first-step kernel.
template <typename T>
int amg_first_launch(const FirstIter<T>& f, long long n, const T* b, const int* skip, hipStream_t s);

// Coarsest-level inverse in the cycle's arithmetic type.
template <typename T> const T* coarse_inv(const Amg& g);
template <> inline const double* coarse_inv<double>(const Amg& g) { return g.inv; }
template <> inline const float* coarse_inv<float>(const Amg& g) { return g.inv32; }
template <typename T> const T* coarse_tiles(const Amg& g);
template <> inline const double* coarse_tiles<double>(const Amg& g) { return g.tiles64; }
template <> inline const float* coarse_tiles<float>(const Amg& g) { return g.tiles32; }

}  // namespace wn
