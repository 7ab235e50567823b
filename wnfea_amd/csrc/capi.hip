// C-ABI of the gfx950 solve-path library (libwnfea_hip.so).
//
// Conventions: every entry point returns int status (0 ok, <0 argument error,
// >0 HIP error code); device arrays are plain device pointers (torch tensors'
// data_ptr() on the Python side); `stream` is a hipStream_t passed as void*
// (null = default stream).  Host-side setup functions take host pointers and
// say so.  Handles own their device buffers; one handle must not be used from
// two threads at once.  Declarations and the reference interface each entry
// replaces are in include/wnfea_hip.h.
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>
#include "wn.hpp"

namespace wn {
// setup.cpp
void rcm_order(int n, const int* ptr, const int* ind, const int* seed_order, int* perm);
long long node_graph(int n_nodes, int M, int npe, const int* conn, const unsigned char* node_free, int* rowptr,
                     int* colidx, long long cap);
long long contributions(int n_nodes, int M, int npe, const int* conn, const int* rowptr, const int* colidx,
                        long long* cptr, int* codes, long long cap);
void node_incidence(int n_nodes, int M, int npe, const int* conn, int* ptr, int* codes);
long long sa_aggregate(int n, const int* ptr, const int* ind, int* agg);
long long bsr_spgemm(int nr, int ra, int ca, int cb, int nc, const int* a_ptr, const int* a_idx, const double* a_val,
                     const int* b_ptr, const int* b_idx, const double* b_val, int* c_ptr, int* c_idx, double* c_val,
                     int threads);
void bsr_transpose(int nr, int nc, int ra, int ca, const int* a_ptr, const int* a_idx, const double* a_val,
                   int* t_ptr, int* t_idx, double* t_val);
long long bsr_symmetrize(int n, int b, const int* ptr, const int* idx, double* val, int threads);
void bsr_spmv_host(int n, int b, const int* ptr, const int* idx, const double* val, const double* x, double* y,
                   int threads);
int sa_tentative(int n_nodes, int b, int k, const int* agg, int n_agg, const double* B, int* p_rowptr, int* p_colidx,
                 double* p_vals, double* Bc);
// solver.hip
struct Amg;
struct AmgLevel;
struct Solver;
int amg_apply(Amg& g, int vt, const void* r, void* z, const int* skip, hipStream_t s, const RhoReq* rq,
              bool first_done);
int amg_level_cheb(Amg& g, int l, int vt, int reps, hipStream_t s);
int amg_set_coarse_fp64(Amg& g, int fp64);
template <typename T>
int pcg_run(Solver& S, const SellMatrix& A, int pkind, const T* dinv, Amg* amg, int amg_vt, const double* scale,
            const T* b, T* x, double rtol, int max_iter, int check_every, const SpmvTune& tune0, hipStream_t s,
            int* iters, double* rel, const MfOp* mf = nullptr);
int ir_run(Solver& S, SellMatrix& A, const double* scale, int pkind, const float* dinv32, Amg* amg, int amg_vt,
           const double* f, double* u, double tol, double inner_rtol, int max_outer, int max_inner, int check_every,
           const SpmvTune& tune, hipStream_t s, int* outer_out, int* inner_out, double* rel_out, int zero_init);
Solver* solver_new();
void solver_delete(Solver*);
Workspace& solver_ws(Solver&);
int pcg_multi(Solver& S, const SellMatrix& A, Amg* g, int vt, const double* scale, int nr, const double* B, double* X,
              double rtol, int max_iter, int check_every, const SpmvTune& t, hipStream_t s, int* iters, double* rel);
int bsr_spmm(const SellMatrix& A, int prec, int nr, const void* X, void* Y, const SpmvTune& t, hipStream_t s);
void solver_set_graphs(Solver&, int enable);
Amg* amg_new();
void amg_delete(Amg*);
int amg_resize(Amg&, int n_levels, int store);
int amg_set_level(Amg& g, int lvl, SellMatrix* fine, int br, int n_rows, const int* a_rowptr, const int* a_colidx,
                  const double* a_vals, const double* dinv, double lmax, double ratio, int degree, int bc_coarse,
                  int n_coarse_rows, const int* p_rowptr, const int* p_colidx, const double* p_vals,
                  const int* r_rowptr, const int* r_colidx, const double* r_vals, int sigma);
int amg_set_coarse(Amg& g, int n, const double* inv);
int amg_set_tune(Amg& g, int prec, int waves, int xcd_remap, int unroll);
int amg_attach_mf(Amg& g, const MfOp* op, const double* scale, hipStream_t s);
int amg_finalize(Amg& g, const SpmvTune& t);
}  // namespace wn

using namespace wn;

struct WnBsr {
  SellMatrix m;
  SpmvTune tune;
  std::unique_ptr<Solver, void (*)(Solver*)> scratch{nullptr, solver_delete};
};

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int wnfea_abi_version(void) { return 1; }

// HBM streaming-read probe: out[grid * 512] per-thread sums of a[bytes / 4]
int wnfea_hbm_stream_read(const float* a, long long bytes, float* out, int grid, void* stream) {
  return launch_stream_read(a, bytes, out, grid, S(stream));
}

// out[0] CUs, out[1] wavefront size, out[2] gfx arch number (e.g. 950),
// out[3] L2 bytes, out[4] LDS bytes per block, out[5] runtime version.
int wnfea_device_info(long long* out) {
  if (!out) return kBadArg;
  int dev = 0;
  WN_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  WN_HIP(hipGetDeviceProperties(&p, dev));
  int rt = 0;
  WN_HIP(hipRuntimeGetVersion(&rt));
  out[0] = p.multiProcessorCount;
  out[1] = p.warpSize;
  int arch = 0;
  if (std::strncmp(p.gcnArchName, "gfx", 3) == 0) arch = static_cast<int>(std::strtol(p.gcnArchName + 3, nullptr, 16));
  out[2] = arch;
  out[3] = p.l2CacheSize;
  out[4] = static_cast<long long>(p.sharedMemPerBlock);
  out[5] = rt;
  return kOk;
}

// ------------------------------------------------------------ element kernels
int wnfea_c3d10_ke(int M, const int* conn, const double* nodes, const double* E, const double* nu, double* Ke,
                   int* bad, void* stream) {
  if (M < 0 || (M > 0 && (!conn || !nodes || !E || !nu || !Ke || !bad))) return kBadArg;
  return launch_ke(M, conn, nodes, E, nu, Ke, bad, S(stream));
}

// Ke kernel of wnfea_c3d10_ke for the process: 0 vector ALU, 1 matrix cores
int wnfea_c3d10_ke_variant(int variant) { return set_ke_variant(variant); }

int wnfea_c3d10_fint(int M, int n_nodes, const int* conn, const double* nodes, const double* E, const double* nu,
                     const double* U, int stride, const int* n2e_ptr, const int* n2e, double* fe_work, double* F,
                     int accumulate, void* stream) {
  if (M < 0 || n_nodes < 0 || (stride != 3 && stride != 6)) return kBadArg;
  if (M > 0 && (!conn || !nodes || !E || !nu || !U || !n2e_ptr || !n2e || !fe_work || !F)) return kBadArg;
  int r = launch_fint_elem(M, conn, nodes, E, nu, U, stride, fe_work, S(stream));
  if (r) return r;
  return launch_gather_node(n_nodes, n2e_ptr, n2e, fe_work, F, stride, accumulate, S(stream));
}

int wnfea_c3d10_stress(int M, int n_nodes, const int* conn, const double* nodes, const double* E, const double* nu,
                       const double* U, const int* n2e_ptr, const int* n2e, double* gp_sig, double* gp_vm,
                       double* cell_sig, double* cell_vm, double* nodal_vm, int* bad, void* stream) {
  if (M < 0 || n_nodes < 0) return kBadArg;
  if (M > 0 && (!conn || !nodes || !E || !nu || !U || !cell_sig || !cell_vm || !bad)) return kBadArg;
  int r = launch_stress(M, conn, nodes, E, nu, U, gp_sig, gp_vm, cell_sig, cell_vm, bad, S(stream));
  if (r) return r;
  if (nodal_vm) {
    if (!n2e_ptr || !n2e) return kBadArg;
    return launch_nodal_average(n_nodes, n2e_ptr, n2e, cell_vm, nodal_vm, S(stream));
  }
  return kOk;
}

int wnfea_adjoint_grad(int M, const int* conn, const double* nodes, const double* E, const double* nu,
                       const double* u, const double* lam, double* fe_work, double* grad, void* stream) {
  if (M < 0 || (M > 0 && (!conn || !nodes || !E || !nu || !u || !lam || !fe_work || !grad))) return kBadArg;
  int r = launch_fint_elem(M, conn, nodes, E, nu, u, 3, fe_work, S(stream));
  if (r) return r;
  return launch_adjoint_dot(M, conn, E, fe_work, lam, grad, S(stream));
}

// ------------------------------------------------------------ beams
int wnfea_beam_forces(int M, int n_nodes, const int* conn, const double* nodes, const double* props, const double* U,
                      const int* n2e_ptr, const int* n2e, double* fe_work, double* F, int accumulate, void* stream) {
  if (M < 0 || n_nodes < 0) return kBadArg;
  if (M > 0 && (!conn || !nodes || !props || !U || !fe_work)) return kBadArg;
  if (n_nodes > 0 && (!n2e_ptr || !n2e || !F)) return kBadArg;
  return launch_beam_forces(M, n_nodes, conn, nodes, props, U, n2e_ptr, n2e, fe_work, F, accumulate, S(stream));
}

int wnfea_beam_tangent(int M, const int* conn, const double* nodes, const double* props, const double* U,
                       double h_rel, double* Kt, void* stream) {
  if (M < 0 || h_rel != h_rel || (M > 0 && (!conn || !nodes || !props || !U || !Kt))) return kBadArg;
  return launch_beam_tangent(M, conn, nodes, props, U, h_rel, Kt, S(stream));
}

int wnfea_beam_ke_linear(int M, const int* conn, const double* nodes, const double* props, double* Ke,
                         void* stream) {
  if (M < 0 || (M > 0 && (!conn || !nodes || !props || !Ke))) return kBadArg;
  return launch_beam_ke_linear(M, conn, nodes, props, Ke, S(stream));
}

// ------------------------------------------------------------ host setup
int wnfea_rcm(int n, const int* ptr, const int* ind, const int* seed_order, int* perm) {
  if (n < 0 || (n > 0 && (!ptr || !ind || !seed_order || !perm))) return kBadArg;
  rcm_order(n, ptr, ind, seed_order, perm);
  return kOk;
}

long long wnfea_node_graph(int n_nodes, int M, int npe, const int* conn, const unsigned char* node_free, int* rowptr,
                           int* colidx, long long cap) {
  if (n_nodes < 0 || M < 0 || npe <= 0 || !rowptr || (M > 0 && !conn)) return kBadArg;
  return node_graph(n_nodes, M, npe, conn, node_free, rowptr, colidx, cap);
}

long long wnfea_contributions(int n_nodes, int M, int npe, const int* conn, const int* rowptr, const int* colidx,
                              long long* cptr, int* codes, long long cap) {
  if (n_nodes < 0 || M < 0 || npe <= 0 || !rowptr || !cptr || (M > 0 && (!conn || !codes))) return kBadArg;
  if ((long long)M * npe * npe > 2147483647LL) return kBadSize;
  return contributions(n_nodes, M, npe, conn, rowptr, colidx, cptr, codes, cap);
}

int wnfea_node_incidence(int n_nodes, int M, int npe, const int* conn, int* ptr, int* codes) {
  if (n_nodes < 0 || M < 0 || npe <= 0 || !ptr || (M > 0 && (!conn || !codes))) return kBadArg;
  node_incidence(n_nodes, M, npe, conn, ptr, codes);
  return kOk;
}

long long wnfea_sa_aggregate(int n, const int* ptr, const int* ind, int* agg) {
  if (n < 0 || (n > 0 && (!ptr || !agg))) return kBadArg;
  return sa_aggregate(n, ptr, ind, agg);
}

int wnfea_sa_tentative(int n_nodes, int b, int k, const int* agg, int n_agg, const double* B, int* p_rowptr,
                       int* p_colidx, double* p_vals, double* Bc) {
  if (n_nodes < 0 || b <= 0 || k <= 0 || n_agg < 0 || !agg || !B || !p_rowptr || !p_colidx || !p_vals || !Bc)
    return kBadArg;
  return sa_tentative(n_nodes, b, k, agg, n_agg, B, p_rowptr, p_colidx, p_vals, Bc);
}

// host block-sparse kernels of the smoothed-aggregation set-up (host pointers)
long long wnfea_host_bsr_spgemm(int nr, int ra, int ca, int cb, int nc, const int* a_ptr, const int* a_idx,
                                const double* a_val, const int* b_ptr, const int* b_idx, const double* b_val,
                                int* c_ptr, int* c_idx, double* c_val, int threads) {
  if (nr < 0 || nc < 0 || ra <= 0 || ca <= 0 || cb <= 0 || !a_ptr || !b_ptr || !c_ptr) return kBadArg;
  if (c_idx && (!c_val || !a_val || !b_val)) return kBadArg;
  return bsr_spgemm(nr, ra, ca, cb, nc, a_ptr, a_idx, a_val, b_ptr, b_idx, b_val, c_ptr, c_idx, c_val, threads);
}

int wnfea_host_bsr_transpose(int nr, int nc, int ra, int ca, const int* a_ptr, const int* a_idx,
                             const double* a_val, int* t_ptr, int* t_idx, double* t_val) {
  if (nr < 0 || nc < 0 || ra <= 0 || ca <= 0 || !a_ptr || !t_ptr) return kBadArg;
  bsr_transpose(nr, nc, ra, ca, a_ptr, a_idx, a_val, t_ptr, t_idx, t_val);
  return kOk;
}

long long wnfea_host_bsr_symmetrize(int n, int b, const int* ptr, const int* idx, double* val, int threads) {
  if (n < 0 || b <= 0 || !ptr) return kBadArg;
  const long long r = bsr_symmetrize(n, b, ptr, idx, val, threads);
  return r < 0 ? kBadArg : r;
}

int wnfea_host_bsr_spmv(int n, int b, const int* ptr, const int* idx, const double* val, const double* x, double* y,
                        int threads) {
  if (n < 0 || b <= 0 || b > 8 || !ptr || !x || !y) return kBadArg;
  bsr_spmv_host(n, b, ptr, idx, val, x, y, threads);
  return kOk;
}

// host set-up of the matrix-free operator's patch layout (host pointers)
int wnfea_host_morton_order(int M, const int* conn, const double* nodes, int* order) {
  if (M < 0 || (M > 0 && (!conn || !nodes || !order))) return kBadArg;
  morton_order(M, conn, nodes, order);
  return kOk;
}

long long wnfea_host_mf_patch_sizes(int M, int n_nodes, const int* conn, const int* order, int elems, int* lmax) {
  if (M < 0 || n_nodes < 0 || elems <= 0 || !lmax || (M > 0 && (!conn || !order))) return kBadArg;
  const long long r = mf_patch_sizes(M, n_nodes, conn, order, elems, lmax);
  return r < 0 ? kBadArg : r;
}

long long wnfea_host_mf_patches(int M, int n_nodes, const int* conn, const int* order, int elems, int stride,
                                int* pnode, unsigned short* pconn, unsigned short* poff, unsigned short* pslot,
                                int* nhead, int* novf) {
  if (M < 0 || n_nodes < 0 || elems <= 0 || !nhead ||
      (M > 0 && (!conn || !order || !pnode || !pconn || !poff || !pslot || !novf)))
    return kBadArg;
  const long long r = mf_patches(M, n_nodes, conn, order, elems, stride, pnode, pconn, poff, pslot, nhead, novf);
  return r < 0 ? kBadArg : r;
}

// ------------------------------------------------------------ block matrices
void* wnfea_bsr_create(int br, int bc, int n_rows, int n_cols, const int* rowptr, const int* colidx, int sigma,
                       int lanes) {
  auto* h = new WnBsr();
  if (h->m.build(br, bc, n_rows, n_cols, rowptr, colidx, sigma, lanes) != kOk) {
    h->m.release();
    delete h;
    return nullptr;
  }
  return h;
}

void wnfea_bsr_destroy(void* h) {
  if (!h) return;
  auto* b = static_cast<WnBsr*>(h);
  b->scratch.reset();
  b->m.release();
  delete b;
}

// out: n_slices, slots, nnzb, sigma, has_v64, has_v32, has_v64e, lanes per row
int wnfea_bsr_info(void* h, long long* out) {
  if (!h || !out) return kBadArg;
  const SellMatrix& m = static_cast<WnBsr*>(h)->m;
  out[0] = m.n_slices;
  out[1] = m.slots;
  out[2] = m.nnzb;
  out[3] = m.sigma;
  out[4] = m.d_v64 != nullptr;
  out[5] = m.d_v32 != nullptr;
  out[6] = m.d_v64e != nullptr;
  out[7] = m.lanes;
  out[8] = m.d_v16 != nullptr;
  return kOk;
}

static bool shape_ok(int waves, int unroll) {
  return (waves == 1 || waves == 2 || waves == 4) && (unroll == 1 || unroll == 2 || unroll == 4);
}

// launch shape of every value precision
int wnfea_bsr_tune(void* h, int waves, int xcd_remap, int unroll) {
  if (!h || !shape_ok(waves, unroll)) return kBadArg;
  static_cast<WnBsr*>(h)->tune.set(-1, waves, xcd_remap ? 1 : 0, unroll);
  return kOk;
}

// launch shape of one value precision: prec 0 FP64 values, 1 FP32 and FP16 values
int wnfea_bsr_tune_precision(void* h, int prec, int waves, int xcd_remap, int unroll) {
  if (!h || !shape_ok(waves, unroll) || (prec != 0 && prec != 1)) return kBadArg;
  static_cast<WnBsr*>(h)->tune.set(prec, waves, xcd_remap ? 1 : 0, unroll);
  return kOk;
}

// which: bit0 FP64 values, bit1 FP32 values (both from the same FP64 input)
int wnfea_bsr_set_values(void* h, const double* vals_std, int which, void* stream) {
  if (!h || (!vals_std && static_cast<WnBsr*>(h)->m.nnzb > 0) || !(which & 3)) return kBadArg;
  return sell_from_std(static_cast<WnBsr*>(h)->m, vals_std, which & 1, which & 2, S(stream));
}

int wnfea_bsr_get_values(void* h, double* vals_std, void* stream) {
  if (!h || !vals_std) return kBadArg;
  return sell_to_std(static_cast<WnBsr*>(h)->m, vals_std, S(stream));
}

int wnfea_bsr_assemble(void* h, const int* rowidx, const int* colidx, const long long* cptr, const int* codes,
                       const double* Ke, const unsigned char* node_fix, double* vals_std, void* stream) {
  if (!h) return kBadArg;
  auto* b = static_cast<WnBsr*>(h);
  if (b->m.nnzb > 0 && (!rowidx || !colidx || !cptr || !codes || !Ke)) return kBadArg;
  return assemble_gather(b->m, rowidx, colidx, cptr, codes, Ke, node_fix, vals_std, S(stream));
}

int wnfea_bsr_assemble_elem(void* h, int npe, int sym, const int* rowidx, const int* colidx, const long long* cptr,
                            const int* codes, const double* Ke, const unsigned char* node_fix, double* vals_std,
                            void* stream) {
  if (!h || !rowidx || !colidx || !cptr || !codes || !Ke) return kBadArg;
  return assemble_elem(static_cast<WnBsr*>(h)->m, npe, sym, rowidx, colidx, cptr, codes, Ke, node_fix, vals_std,
                       S(stream));
}

int wnfea_bsr_diag(void* h, const long long* diag_block, double* diag, void* stream) {
  if (!h || !diag_block || !diag) return kBadArg;
  return extract_diag(static_cast<WnBsr*>(h)->m, diag_block, diag, S(stream));
}

int wnfea_bsr_equilibrate(void* h, const int* rowidx, const int* colidx, const double* scale, int want64e,
                          int want16, void* stream) {
  if (!h || !rowidx || !colidx || !scale) return kBadArg;
  return equilibrate(static_cast<WnBsr*>(h)->m, rowidx, colidx, scale, want64e, S(stream), want16);
}

// prec: 0 = FP64 primary values, 1 = FP32 values, 2 = FP64 equilibrated values
int wnfea_bsr_spmv(void* h, int prec, const void* x, void* y, double alpha, double beta, void* stream) {
  if (!h || !x || !y) return kBadArg;
  auto* b = static_cast<WnBsr*>(h);
  switch (prec) {
    case 0:
      return spmv_axpby<double, double>(b->m, static_cast<const double*>(x), static_cast<double*>(y), alpha, beta,
                                        b->tune, S(stream));
    case 1:
      return spmv_axpby<float, float>(b->m, static_cast<const float*>(x), static_cast<float*>(y), alpha, beta,
                                      b->tune, S(stream));
    case 2:
      if (!b->m.d_v64e) return kNotReady;
      return spmv_axpby<double, double>(b->m.equilibrated(), static_cast<const double*>(x), static_cast<double*>(y),
                                        alpha, beta, b->tune, S(stream));
    default:
      return kBadMode;
  }
}

// y = A x and out[0] = x . (A x), FP64 (prec 0/2) or FP32 values with FP64 dot (prec 1)
int wnfea_bsr_spmv_dot(void* h, int prec, const void* x, void* y, double* out, void* stream) {
  if (!h || !x || !y || !out) return kBadArg;
  auto* b = static_cast<WnBsr*>(h);
  if (!b->scratch) b->scratch.reset(solver_new());
  Workspace& w = solver_ws(*b->scratch);
  int r;
  if (prec == 1)
    r = spmv_dot<float>(b->m, static_cast<const float*>(x), static_cast<float*>(y), w, 0, b->tune, S(stream));
  else if (prec == 0 || prec == 2)
    r = spmv_dot<double>(prec == 2 ? b->m.equilibrated() : b->m, static_cast<const double*>(x),
                         static_cast<double*>(y), w, 0, b->tune, S(stream));
  else
    return kBadMode;
  if (r) return r;
  WN_HIP(hipMemcpyAsync(out, w.sc + kScGamma, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
  WN_HIP(hipStreamSynchronize(S(stream)));
  return kOk;
}

// ------------------------------------------------------------ AMG
// store: bit0 FP64 values, bit1 FP32 values, bit2 FP16 values (with FP32 arithmetic)
void* wnfea_amg_create(int n_levels, int store) {
  if (n_levels <= 0 || n_levels > 32 || !(store & 7)) return nullptr;
  Amg* g = amg_new();
  if (amg_resize(*g, n_levels, store) != kOk) {
    amg_delete(g);
    return nullptr;
  }
  return g;
}

void wnfea_amg_destroy(void* h) {
  if (h) amg_delete(static_cast<Amg*>(h));
}

int wnfea_amg_set_level(void* h, int lvl, void* fine_bsr, int br, int n_rows, const int* a_rowptr,
                        const int* a_colidx, const double* a_vals, const double* dinv, double lmax, double ratio,
                        int degree, int bc_coarse, int n_coarse_rows, const int* p_rowptr, const int* p_colidx,
                        const double* p_vals, const int* r_rowptr, const int* r_colidx, const double* r_vals,
                        int sigma) {
  if (!h || !dinv) return kBadArg;
  SellMatrix* fine = fine_bsr ? &static_cast<WnBsr*>(fine_bsr)->m : nullptr;
  return amg_set_level(*static_cast<Amg*>(h), lvl, fine, br, n_rows, a_rowptr, a_colidx, a_vals, dinv, lmax, ratio,
                       degree, bc_coarse, n_coarse_rows, p_rowptr, p_colidx, p_vals, r_rowptr, r_colidx, r_vals,
                       sigma);
}

int wnfea_amg_set_coarse(void* h, int n, const double* inv) {
  if (!h || n <= 0 || !inv) return kBadArg;
  return amg_set_coarse(*static_cast<Amg*>(h), n, inv);
}

// allocate the cycle vectors of the stored precisions; set the SpMV launch shape
int wnfea_amg_finalize(void* h, int waves, int xcd_remap, int unroll) {
  if (!h || !shape_ok(waves, unroll)) return kBadArg;
  SpmvTune t;
  t.set(-1, waves, xcd_remap ? 1 : 0, unroll);
  return amg_finalize(*static_cast<Amg*>(h), t);
}

// SpMV launch shape of the cycle for one value precision (prec 0 FP64 values,
// 1 FP32 and FP16 values), after wnfea_amg_finalize
int wnfea_amg_tune_precision(void* h, int prec, int waves, int xcd_remap, int unroll) {
  if (!h || !shape_ok(waves, unroll) || (prec != 0 && prec != 1)) return kBadArg;
  return amg_set_tune(*static_cast<Amg*>(h), prec, waves, xcd_remap ? 1 : 0, unroll);
}

// z = M r (one V-cycle); vt 0 = FP64, 1 = FP32, 2 = FP16 values, 3 = FP16 on the
// finest level and FP32 below, 4 = FP32 finest and FP16 below (vt >= 1: FP32 vectors)
int wnfea_amg_vcycle(void* h, int vt, const void* r, void* z, void* stream) {
  if (!h || !r || !z) return kBadArg;
  return amg_apply(*static_cast<Amg*>(h), vt, r, z, nullptr, S(stream), nullptr, false);
}

// FP32 cycles: coarsest solve with the FP64 inverse (fp64 != 0) or its FP32 copy
int wnfea_amg_set_coarse_fp64(void* h, int fp64) {
  if (!h) return kBadArg;
  return amg_set_coarse_fp64(*static_cast<Amg*>(h), fp64);
}

// reps Chebyshev smoothing steps of one level on its own vectors (vt 0 FP64,
// 1 FP32 values): the cycle's smoother kernel alone, for timing
int wnfea_amg_level_cheb(void* h, int lvl, int vt, int reps, void* stream) {
  if (!h) return kBadArg;
  return amg_level_cheb(*static_cast<Amg*>(h), lvl, vt, reps, S(stream));
}

// ------------------------------------------------------------ solvers
void* wnfea_solver_create(void) { return solver_new(); }
void wnfea_solver_destroy(void* h) {
  if (h) solver_delete(static_cast<Solver*>(h));
}

// enable != 0: run full chunks of PCG iterations as a replayed HIP graph
int wnfea_solver_set_graphs(void* h, int enable) {
  if (!h) return kBadArg;
  solver_set_graphs(*static_cast<Solver*>(h), enable);
  return kOk;
}

// PCG from x = 0 until ||r|| <= rtol ||b|| (recursive residual).
// prec: 0 FP64 primary values, 1 FP32 values (FP32 vectors, FP64 dots),
// 2 FP64 equilibrated values.  pkind: 0 none, 1 Jacobi (dinv, same precision
// as the vectors), 2 AMG (amg handle; amg_vt 0 FP64 / 1 FP32 / 2 FP16 /
// 3 FP16-finest-level cycle / 4 FP16-below-finest cycle / 5 FP16 smoothing
// steps, the nonsymmetric cycle run with the flexible beta;
// scale, when given, makes the preconditioner z = S M(S r)).
int wnfea_pcg(void* solver, void* bsr, int prec, int pkind, const void* dinv, void* amg, int amg_vt,
              const double* scale, const void* b, void* x, double rtol, int max_iter, int check_every, int* iters,
              double* rel, void* stream) {
  if (!solver || !bsr || !b || !x || max_iter < 0 || pkind < 0 || pkind > 2 || amg_vt < 0 || amg_vt > 5)
    return kBadArg;
  if ((pkind == 1 && !dinv) || (pkind == 2 && !amg)) return kBadArg;
  auto* B = static_cast<WnBsr*>(bsr);
  Solver& Sv = *static_cast<Solver*>(solver);
  Amg* g = static_cast<Amg*>(amg);
  if (prec == 1)
    return pcg_run<float>(Sv, B->m, pkind, static_cast<const float*>(dinv), g, amg_vt, scale,
                          static_cast<const float*>(b), static_cast<float*>(x), rtol, max_iter, check_every, B->tune,
                          S(stream), iters, rel);
  if (prec == 0 || prec == 2) {
    if (prec == 2 && !B->m.d_v64e) return kNotReady;
    return pcg_run<double>(Sv, prec == 2 ? B->m.equilibrated() : B->m, pkind, static_cast<const double*>(dinv), g,
                           amg_vt, scale, static_cast<const double*>(b), static_cast<double*>(x), rtol, max_iter,
                           check_every, B->tune, S(stream), iters, rel);
  }
  return kBadMode;
}

// FP64 iterative refinement with an FP32 inner PCG (see solver.hip); amg_vt 1 or 2.
int wnfea_ir(void* solver, void* bsr, const double* scale, int pkind, const float* dinv32, void* amg, int amg_vt,
             const double* f, double* u, double tol, double inner_rtol, int max_outer, int max_inner,
             int check_every, int zero_init, int* outer, int* inner, double* rel, void* stream) {
  if (!solver || !bsr || !scale || !f || !u || pkind < 0 || pkind > 2) return kBadArg;
  if ((pkind == 1 && !dinv32) || (pkind == 2 && (!amg || amg_vt < 1 || amg_vt > 4))) return kBadArg;
  auto* B = static_cast<WnBsr*>(bsr);
  return ir_run(*static_cast<Solver*>(solver), B->m, scale, pkind, dinv32, static_cast<Amg*>(amg), amg_vt, f, u, tol,
                inner_rtol, max_outer, max_inner, check_every, B->tune, S(stream), outer, inner, rel, zero_init);
}

// PCG on nrhs (4 or 8) interleaved columns B[n][nrhs] -> X[n][nrhs], FP64
// vectors and SpMM on the primary values, FP32 V-cycle (amg_vt 1) on S K S.
int wnfea_pcg_multi(void* solver, void* bsr, void* amg, int amg_vt, const double* scale, int nrhs, const double* B,
                    double* X, double rtol, int max_iter, int check_every, int* iters, double* rel, void* stream) {
  if (!solver || !bsr || !amg || !B || !X || max_iter < 0 || (nrhs != 4 && nrhs != 8)) return kBadArg;
  auto* Bm = static_cast<WnBsr*>(bsr);
  if (!Bm->m.d_v64) return kNotReady;
  Solver& Sv = *static_cast<Solver*>(solver);
  return pcg_multi(Sv, Bm->m, static_cast<Amg*>(amg), amg_vt, scale, nrhs, B, X, rtol, max_iter, check_every,
                   Bm->tune, S(stream), iters, rel);
}

// Y = K X on nrhs (4 or 8) interleaved columns; prec 0 FP64, 1 FP32 values.
int wnfea_bsr_spmm(void* h, int prec, int nrhs, const void* X, void* Y, void* stream) {
  if (!h || !X || !Y) return kBadArg;
  auto* b = static_cast<WnBsr*>(h);
  return bsr_spmm(b->m, prec, nrhs, X, Y, b->tune, S(stream));
}

// ---------------------------------------------------------------- matrix-free operator
struct WnMf {
  MfOp op;
};

int wnfea_mf_create(int M, int n_nodes, const int* conn, const double* nodes, const double* E, const double* nu,
                    const unsigned char* node_fix, const int* n2e_ptr, const int* n2e, int* bad, void** out,
                    void* stream) {
  if (!out) return kBadArg;
  *out = nullptr;
  if (M < 0 || n_nodes <= 0 || !n2e_ptr || (M > 0 && (!conn || !nodes || !E || !nu || !n2e))) return kBadArg;
  auto* h = new WnMf();
  MfOp& op = h->op;
  op.M = M;
  op.n_nodes = n_nodes;
  op.conn = conn;
  op.E = E;
  op.nu = nu;
  op.fix = node_fix;
  op.n2e_ptr = n2e_ptr;
  op.n2e = n2e;
  const int r = mf_setup(op, nodes, bad, S(stream));
  if (r) {
    mf_release(op);
    delete h;
    return r;
  }
  *out = h;
  return kOk;
}

void wnfea_mf_destroy(void* h) {
  if (!h) return;
  auto* m = static_cast<WnMf*>(h);
  mf_release(m->op);
  delete m;
}

// 1 when the mesh is affine (one set of quadrature data per element), 0 when
// it holds curved elements (data per Gauss point)
int wnfea_mf_affine(void* h) {
  if (!h) return kBadArg;
  return static_cast<WnMf*>(h)->op.affine ? 1 : 0;
}

// Patch layout of the operator's products: elems > 0 (a multiple of 64, at
// most 256) rebuilds it with that many elements per patch, 0 drops it (the
// per-element contributions and node gather are used instead).  Affine
// meshes only (-3 otherwise); -2 when one patch of the FP64 product would
// not fit a workgroup's LDS.  Hierarchies attached before keep running, on
// the per-element path, until attached again.
int wnfea_mf_set_patch(void* h, int elems, void* stream) {
  if (!h) return kBadArg;
  return mf_set_patch(static_cast<WnMf*>(h)->op, elems, S(stream));
}

// out[0..5] = elements per patch (0: no patch layout), patches, most local
// nodes of one patch, local nodes of all patches, affine (1/0), the
// per-patch stride of the local-node arrays
int wnfea_mf_info(void* h, long long* out) {
  if (!h || !out) return kBadArg;
  const MfOp& op = static_cast<WnMf*>(h)->op;
  out[0] = op.patch;
  out[1] = op.n_patch;
  out[2] = op.lmax;
  out[3] = op.n_local;
  out[4] = op.affine;
  out[5] = op.stride;
  return kOk;
}


// y = K u (device FP64 vectors of 3 * n_nodes)
int wnfea_mf_apply(void* h, const double* u, double* y, void* stream) {
  if (!h || !u || !y) return kBadArg;
  return mf_apply(static_cast<WnMf*>(h)->op, u, nullptr, y, nullptr, 0, nullptr, 0, S(stream));
}

// r = f - K u and out[0] = ||r||, out[1] = ||f|| (as wnfea_residual, matrix-free)
int wnfea_mf_residual(void* solver, void* h, const double* u, const double* f, double* r, double* out, void* stream) {
  if (!solver || !h || !u || !f || !r || !out) return kBadArg;
  const MfOp& op = static_cast<WnMf*>(h)->op;
  Workspace& w = solver_ws(*static_cast<Solver*>(solver));
  const int nb = mf_grid(op);
  int e = w.alloc(nb + 8);
  if (e) return e;
  if ((e = mf_apply(op, u, f, r, w.partials, 2, nullptr, 0, S(stream)))) return e;
  if ((e = launch_finalize(w.partials, nb, 1, kFinStore, w.sc + kScTrue, w.sc, nullptr, 0, nullptr, S(stream))))
    return e;
  if ((e = launch_dot<double>(3LL * op.n_nodes, f, f, w, w.sc + kScF2, S(stream)))) return e;
  double hs[kScCount];
  WN_HIP(hipMemcpyAsync(hs, w.sc, sizeof(double) * kScCount, hipMemcpyDeviceToHost, S(stream)));
  WN_HIP(hipStreamSynchronize(S(stream)));
  out[0] = std::sqrt(hs[kScTrue]);
  out[1] = std::sqrt(hs[kScF2]);
  return kOk;
}

// Run the finest level's smoothing and defect products of the FP32 and FP64
// V-cycles matrix-free on diag(scale) K diag(scale) (scale: the device FP64
// equilibration of the fine matrix, 3 * n_nodes); mf = null detaches.  The
// operator and the scale must outlive the attachment.
int wnfea_amg_attach_mf(void* amg, void* mf, const double* scale, void* stream) {
  if (!amg || (mf && !scale)) return kBadArg;
  return amg_attach_mf(*static_cast<Amg*>(amg), mf ? &static_cast<WnMf*>(mf)->op : nullptr, scale, S(stream));
}

// FP64 PCG with the matrix-free operator (pkind and preconditioner as in wnfea_pcg)
int wnfea_pcg_mf(void* solver, void* h, int pkind, const double* dinv, void* amg, int amg_vt, const double* scale,
                 const double* b, double* x, double rtol, int max_iter, int check_every, int* iters, double* rel,
                 void* stream) {
  if (!solver || !h || !b || !x || max_iter < 0 || pkind < 0 || pkind > 2 || amg_vt < 0 || amg_vt > 5)
    return kBadArg;
  if ((pkind == 1 && !dinv) || (pkind == 2 && !amg)) return kBadArg;
  static const SellMatrix none;
  SpmvTune t;
  return pcg_run<double>(*static_cast<Solver*>(solver), none, pkind, dinv, static_cast<Amg*>(amg), amg_vt, scale, b,
                         x, rtol, max_iter, check_every, t, S(stream), iters, rel, &static_cast<WnMf*>(h)->op);
}

// r = f - K u (FP64 primary values) and out[0] = ||r||, out[1] = ||f||.
int wnfea_residual(void* solver, void* bsr, const double* u, const double* f, double* r, double* out, void* stream) {
  if (!solver || !bsr || !u || !f || !out) return kBadArg;
  auto* B = static_cast<WnBsr*>(bsr);
  Solver& Sv = *static_cast<Solver*>(solver);
  Workspace& w = solver_ws(Sv);
  int e = w.alloc(B->m.n_slices + 8);
  if (e) return e;
  const long long n = (long long)B->m.n_rows * B->m.br;
  SpmvTune t = B->tune;
  if ((e = spmv_residual64(B->m, u, f, r, nullptr, nullptr, w, w.sc + kScTrue, t, S(stream)))) return e;
  if ((e = launch_dot<double>(n, f, f, w, w.sc + kScF2, S(stream)))) return e;
  double h[kScCount];
  WN_HIP(hipMemcpyAsync(h, w.sc, sizeof(double) * kScCount, hipMemcpyDeviceToHost, S(stream)));
  WN_HIP(hipStreamSynchronize(S(stream)));
  out[0] = std::sqrt(h[kScTrue]);
  out[1] = std::sqrt(h[kScF2]);
  return kOk;
}

}  // extern "C"

// ------------------------------------------------------------ no allocation
// Scratch of wnfea_residual_noalloc: the reduction partials (up to
// 4 * kNoAllocBlocks values, as a Workspace holds them) followed by the
// device scalar slots.  Each GPU the library runs on holds its own copy of
// this array until the process ends.  Host code never passes the symbol
// itself to a kernel: it takes the array's device address on each device
// once (hipGetSymbolAddress) and uses that address.
constexpr int kNoAllocBlocks = 65536;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[4 * kNoAllocBlocks + kScCount];
static double* g_scratch_addr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once per device).
static int scratch_address(double** out) {
  int dev = -1;
  WN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return kBadArg;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    WN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return kNotReady;
    g_scratch_addr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_addr[dev];
  return kOk;
}

__global__ void norms_from_squares_kernel(const double* sc, double* out) {
  if (threadIdx.x == 0) {
    out[0] = sqrt(sc[kScTrue]);
    out[1] = sqrt(sc[kScF2]);
  }
}

extern "C" {

// r = f - K u (when r is not null) and out_dev[0] = ||f - K u||,
// out_dev[1] = ||f|| on the device: wnfea_residual's kernels with their
// reduction partials in the library's static per-GPU scratch instead of a
// solver workspace.  Allocates nothing, copies nothing to the host and does
// not synchronize.  Calls that overlap on one GPU share the scratch and race.
int wnfea_residual_noalloc(void* bsr, const double* u, const double* f, double* r, double* out_dev, void* stream) {
  if (!bsr || !u || !f || !out_dev) return kBadArg;
  auto* B = static_cast<WnBsr*>(bsr);
  const SellMatrix& A = B->m;
  if (A.br != A.bc) return kBadArg;
  // the SpMV's partial blocks (n_slices rounded up to the 8 XCDs) and the dot's
  // (at most 2048) must fit the scratch
  if ((long long)A.n_slices + 8 > kNoAllocBlocks) return kBadSize;
  double* base = nullptr;
  int e = scratch_address(&base);
  if (e) return e;
  Workspace w;  // a view of the scratch: Workspace::alloc sees enough room and allocates nothing
  w.partials = base;
  w.sc = base + 4LL * kNoAllocBlocks;
  w.max_blocks = kNoAllocBlocks;
  const long long n = (long long)A.n_rows * A.br;
  hipStream_t s = S(stream);
  if ((e = launch_fill<double>(kScCount, w.sc, 0.0, s))) return e;  // a kernel, not a memset: capturable
  if ((e = spmv_residual64(A, u, f, r, nullptr, nullptr, w, w.sc + kScTrue, B->tune, s))) return e;
  if ((e = launch_dot<double>(n, f, f, w, w.sc + kScF2, s))) return e;
  hipLaunchKernelGGL(norms_from_squares_kernel, dim3(1), dim3(64), 0, s, w.sc, out_dev);
  return static_cast<int>(hipGetLastError());
}

}  // extern "C"
