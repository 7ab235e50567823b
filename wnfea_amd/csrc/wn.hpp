// Internal declarations shared by the library's translation units.
#pragma once
#include <vector>
#include "common.hpp"
#include "sell.hpp"

namespace wn {

// Reduction scratch of one stream: per-block partials (see store_block_partials)
// and the device scalar slots.
struct Workspace {
  double* partials = nullptr;
  double* sc = nullptr;
  int max_blocks = 0;
  int alloc(int max_blocks_needed);
  void release();
};

// Sliced-ELL block matrix owned by the library (device buffers).
struct SellMatrix {
  int br = 3, bc = 3;
  int n_rows = 0, n_cols = 0, n_slices = 0, sigma = 0, lanes = 1;
  long long nnzb = 0, slots = 0;
  long long* d_off = nullptr;
  int2* d_rinfo = nullptr;  // [n_slices * 64/lanes] {block row or -1, blocks in the row}
  int* d_col = nullptr;
  long long* d_slot = nullptr;  // standard block index -> slot
  double* d_v64 = nullptr;   // primary FP64 values (the assembled K)
  double* d_v64e = nullptr;  // optional FP64 equilibrated values S K S
  float* d_v32 = nullptr;    // FP32 values (equilibrated for the fine level)
  _Float16* d_v16 = nullptr; // optional FP16 copy of the FP32 values (preconditioner storage)
  // host copies of the standard structure (for exports and AMG setup)
  std::vector<int> h_rowptr, h_colidx;

  // lanes_ <= 0 picks the lanes per row from the shape (see sell.hip)
  int build(int br_, int bc_, int n_rows_, int n_cols_, const int* rowptr, const int* colidx, int sigma_,
            int lanes_ = 0);
  int ensure_values(int want64, int want32, int want16 = 0);
  void release();
  // Non-owning alias whose FP64 values are the equilibrated copy (when present).
  SellMatrix* eq_alias = nullptr;
  SellMatrix& equilibrated();
  template <typename T> SellView<T> view() const;
};

template <> inline SellView<double> SellMatrix::view<double>() const {
  return SellView<double>{n_rows, n_slices, lanes, d_off, d_rinfo, d_col, d_v64};
}
template <> inline SellView<float> SellMatrix::view<float>() const {
  return SellView<float>{n_rows, n_slices, lanes, d_off, d_rinfo, d_col, d_v32};
}
template <> inline SellView<_Float16> SellMatrix::view<_Float16>() const {
  return SellView<_Float16>{n_rows, n_slices, lanes, d_off, d_rinfo, d_col, d_v16};
}

// SpMV launch knobs (picked by the autotuner; defaults are sane).
struct SpmvTune {
  // launch shape for 4- and 2-byte values ...
  int waves = 1;               // wavefronts per slice (1, 2, 4)
  int unroll = 1;              // blocks in flight per lane (1, 2, 4; capped by register use)
  int xcd_remap = 1;           // contiguous slice ranges per XCD
  // ... and for 8-byte values (the best shapes differ: more waves per slice
  // for the longer FP64 block stream)
  int waves64 = 1, unroll64 = 1, xcd_remap64 = 1;
  const int* flags = nullptr;  // device loop flags (see guard_skip in sell.hpp)
  int guard = 0;               // 0 always run, 1 skip once the done flag is set
  template <typename VT> int waves_for() const { return sizeof(VT) == 8 ? waves64 : waves; }
  template <typename VT> int unroll_for() const { return sizeof(VT) == 8 ? unroll64 : unroll; }
  template <typename VT> int xcd_for() const { return sizeof(VT) == 8 ? xcd_remap64 : xcd_remap; }
  void set(int prec, int w, int x, int u) {  // prec 0: 8-byte values, 1: 4/2-byte values, -1: both
    if (prec <= 0) { waves64 = w; xcd_remap64 = x; unroll64 = u; }
    if (prec != 0) { waves = w; xcd_remap = x; unroll = u; }
  }
};

// -------- element kernels (element.hip)
int upload_element_consts(hipStream_t s);
int set_ke_variant(int v);
int launch_ke(int M, const int* conn, const double* nodes, const double* E, const double* nu, double* Ke, int* bad,
              hipStream_t s);
// Matrix-free C3D10 operator with the assembled matrix's Dirichlet rule.
// Mesh arrays are borrowed (device); quadrature data and the element work
// buffer are owned.
struct MfOp {
  int M = 0, n_nodes = 0;
  const int* conn = nullptr;
  const double* E = nullptr;
  const double* nu = nullptr;
  const unsigned char* fix = nullptr;  // per node, bit c: DOF c fixed (may be null)
  const int* n2e_ptr = nullptr;
  const int* n2e = nullptr;
  // inverse Jacobian entries and det/24: [10][M][4] per Gauss point, or
  // [10][M] once per element when every element is affine (straight edges,
  // midside nodes at the edge midpoints: J is then the same at every point)
  double* qd = nullptr;
  int affine = 0;
  double* fe = nullptr;    // [10][M][3] element contributions per local node
  double* mask = nullptr;  // [3 n_nodes] 0 for a fixed DOF, else 1
  int2* chunk = nullptr;  // [n_chunks+1] {first node, first incidence} of each gather chunk
  int n_chunks = 0;
  // Patch layout (affine meshes): the elements sorted along a Morton curve of
  // their centroids and cut into patches of `patch` consecutive slots.  One
  // workgroup per patch stages the patch's nodal values in LDS, keeps its
  // element forces in LDS and writes one partial sum per (patch, local node)
  // to slot p stride + j of the partial array; a node-ordered kernel then
  // adds each node's partials in patch order.  Per-patch arrays have a fixed
  // stride (> the most local nodes of one patch), so a patch's data is found
  // without a header.  patch == 0: no patch layout (per-element
  // contributions in `fe`, gathered by `chunk`).
  int patch = 0, n_patch = 0, lmax = 0, stride = 0;  // elements per patch, patches, most local nodes, stride
  int n_local = 0;                       // local nodes of all patches (= partial sums)
  int gen = 0;                           // layout generation: smoothers built on another one fall back
  int* pe = nullptr;                     // [M] element of each slot
  int* pnode = nullptr;                  // [n_patch][stride] global node of each local node (-1: none)
  unsigned short* pconn = nullptr;       // [M][10] local node of each element corner
  unsigned short* poff = nullptr;        // [n_patch][stride] start of the local node's slot list (patch-relative)
  unsigned short* pslot = nullptr;       // [M][10] per patch: the local nodes' slot lists (slot = 10 e + corner)
  int4* nhead = nullptr;                 // [n_nodes] first three partial slots and the count (or overflow)
  int* novf = nullptr;                   // overflow lists of nodes with more than three partials
  double* qdp = nullptr;                 // [M][10] affine quadrature data, one record per slot
  double* part = nullptr;                // [n_patch * stride][3]
  std::vector<int> order;                // host: elements in Morton order of their centroids
};
// The finest level's smoothing and defect products of a V-cycle run
// matrix-free on the equilibrated operator: y = s .* K (s .* x), with K the
// matrix-free C3D10 operator (assembled Dirichlet rule) and s the diagonal
// scaling.  FP64 uses the operator's own quadrature data and element buffer;
// FP32 keeps rounded copies of the quadrature data and the scaling.
struct MfSmooth {
  const MfOp* op = nullptr;     // not owned: it must outlive the attachment
  const double* s64 = nullptr;  // not owned
  double* s64e = nullptr;       // s with fixed DOFs zeroed (the element kernel's input scaling)
  float* s32 = nullptr;
  float* s32e = nullptr;
  float* qd32 = nullptr;  // rounded copy of the operator's qd (same layout)
  float* fe32 = nullptr;  // [10][M][3]
  float* qdp32 = nullptr;   // [M][12] patch layout: rounded inverse Jacobians and w c11, w c12, w c44
  float* part32 = nullptr;  // [n_patch * stride][3]
  int gen = -1;             // the operator's layout generation the copies were made for
};
int mf_smooth_attach(MfSmooth& m, const MfOp* op, const double* scale, int want32, hipStream_t s);
void mf_smooth_release(MfSmooth& m);
template <typename T> bool mf_smooth_ready(const MfSmooth& m);
template <> bool mf_smooth_ready<double>(const MfSmooth& m);
template <> bool mf_smooth_ready<float>(const MfSmooth& m);
// x_out = x_in + d',  d' = c1 d + c2 dinv (b - A x_in); dinv null: D = I
// (the equilibrated finest level, whose inverse diagonal is 1 everywhere)
template <typename T>
int mf_smooth_cheb(const MfSmooth& m, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, T c1, T c2,
                   const int* skip, hipStream_t s);
// res = b - A x
template <typename T>
int mf_smooth_defect(const MfSmooth& m, const T* x, const T* b, T* res, const int* skip, hipStream_t s);
// Partials of rho = r . (S z) requested from the Chebyshev step that writes
// a V-cycle's output z (scale null: S = I); nb partials, see
// mf_smooth_rho_blocks.
struct RhoReq {
  const double* r;
  const double* scale;
  double* partials;
};
// The Chebyshev step above, also writing the RhoReq partials of its x_out
// (patch layout only: kBadMode otherwise).
template <typename T>
int mf_smooth_cheb_rho(const MfSmooth& m, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, T c1, T c2,
                       const RhoReq& rq, const int* skip, hipStream_t s);
// Partials mf_smooth_cheb_rho writes in precision T (0: it cannot run).
template <typename T>
int mf_smooth_rho_blocks(const MfSmooth& m);
std::vector<int> mf_chunks(const std::vector<int>& n2e_ptr);
int mf_setup(MfOp& op, const double* nodes, int* bad, hipStream_t s);
void mf_release(MfOp& op);
int mf_grid(const MfOp& op);  // blocks of the gather (= partial count of modes 1 and 2)
// Build (elems > 0: elements per patch, a multiple of 64, at most 256) or drop
// (elems == 0) the patch layout; smoothers attached before fall back to the
// per-element path until attached again.
int mf_set_patch(MfOp& op, int elems, hipStream_t s);
// host set-up of the patch layout (setup.cpp)
void morton_order(int M, const int* conn, const double* X, int* order);
long long mf_patch_sizes(int M, int N, const int* conn, const int* order, int P, int* lmax);
long long mf_patches(int M, int N, const int* conn, const int* order, int P, int stride, int* pnode,
                     unsigned short* pconn, unsigned short* poff, unsigned short* pslot, int* nhead, int* novf);
int mf_apply(const MfOp& op, const double* u, const double* f, double* y, double* partials, int mode,
             const int* flags, int guard, hipStream_t s);
int launch_fint_elem(int M, const int* conn, const double* nodes, const double* E, const double* nu, const double* U,
                     int stride, double* fe, hipStream_t s);
int launch_gather_node(int n_nodes, const int* n2e_ptr, const int* n2e, const double* fe, double* out, int stride,
                       int accumulate, hipStream_t s);
int launch_stress(int M, const int* conn, const double* nodes, const double* E, const double* nu, const double* U,
                  double* gp_sig, double* gp_vm, double* cell_sig, double* cell_vm, int* bad, hipStream_t s);
int launch_nodal_average(int n_nodes, const int* n2e_ptr, const int* n2e, const double* cell_vm, double* nodal_vm,
                         hipStream_t s);
int launch_beam_forces(int M, int n_nodes, const int* conn, const double* nodes, const double* props,
                       const double* U, const int* n2e_ptr, const int* n2e, double* fe, double* F, int accumulate,
                       hipStream_t s);
int launch_beam_tangent(int M, const int* conn, const double* nodes, const double* props, const double* U,
                        double h_rel, double* Kt, hipStream_t s);
int launch_beam_ke_linear(int M, const int* conn, const double* nodes, const double* props, double* Ke,
                          hipStream_t s);
int launch_adjoint_dot(int M, const int* conn, const double* E, const double* fe, const double* lam, double* grad,
                       hipStream_t s);

// -------- vector kernels (blas.hip)
int vec_grid(long long n);
// one-block finalisation of block partials (FinMode in common.hpp)
int launch_finalize(const double* partials, int nb, int K, int mode, double* out, double* sc, int* flags, int first,
                    const int* skip, hipStream_t s);
template <typename T> int launch_fill(long long n, T* x, T v, hipStream_t s);
template <typename A, typename B>
int launch_cast_scale(long long n, const A* x, const double* sc, double f, B* y, hipStream_t s);
// fin == false: leave the partials of rho for launch_update_p to finalise
template <typename T>
int launch_jacobi_dot(long long n, const T* r, const T* dinv, T* z, Workspace& w, int first, const int* skip,
                      hipStream_t s, bool fin = true);
// rho = r . (S z) (S may be null); finalise beta = rho / rho_prev, or with q
// (= A p of the previous iteration) the flexible beta = -alpha (S z . q) / rho_prev.
// fin == false (without q): only the partials of rho, for launch_update_p
template <typename T, typename ZT>
int launch_rho(long long n, const T* r, const ZT* z, const double* scale, Workspace& w, int first, const int* skip,
               hipStream_t s, const T* q = nullptr, bool fin = true);
// p = S z + beta p (p = S z when first); with rho_part the kernel finalises
// rho from its nb_rho partials and forms beta itself
template <typename T, typename ZT>
int launch_update_p(long long n, const ZT* z, const double* scale, T* p, Workspace& w, int first, const int* skip,
                    hipStream_t s, const double* rho_part = nullptr, int nb_rho = 0);
// The finest level's first Chebyshev iterate of the next V-cycle, written by
// the kernel that writes the cycle's input b: x1 = d1 = c2 dinv .* b.
template <typename T>
struct FirstIter {
  const T* dinv = nullptr;
  bool unit = false;  // dinv is 1 everywhere: the x/r update does not read it
  T* x1 = nullptr;
  T* d1 = nullptr;
  T c2 = T(0);
};
// x += alpha p; r -= alpha q; rr; optional rp = S r (and with fi the next
// cycle's first iterate from it); raises flags[0] (done) when
// rr <= sc[kScTol2]; with gam_part the kernel finalises gamma = p . A p from
// its nb_gam partials and forms alpha itself
template <typename T, typename RT>
int launch_update_xr(long long n, T* x, T* r, const T* p, const T* q, RT* rp, const double* scale, Workspace& w,
                     int* flags, hipStream_t s, const double* gam_part = nullptr, int nb_gam = 0,
                     const FirstIter<RT>* fi = nullptr);
template <typename T> int launch_dot(long long n, const T* a, const T* b, Workspace& w, double* out, hipStream_t s);
template <typename T>
int launch_ir_update(long long n, double* u, const T* d, const double* sc_vec, const double* nrm2, hipStream_t s);
template <typename T> int launch_normalise(long long n, const T* x, T* y, const double* nrm2, hipStream_t s);
// HBM streaming-read probe (blas.hip): a[bytes / 4] floats, 16-byte aligned,
// bytes a multiple of 16; out[grid * 512] per-thread sums.
int launch_stream_read(const float* a, long long bytes, float* out, int grid, hipStream_t s);

// -------- sliced-ELL kernels (sell.hip)
// y = alpha A x + beta y
template <typename VT, typename XT>
int spmv_axpby(const SellMatrix& A, const XT* x, XT* y, double alpha, double beta, const SpmvTune& t, hipStream_t s);
// y = A p, gamma = p.y (into sc[kScGamma]); alpha = rho/gamma if with_alpha
template <typename VT>
int spmv_dot(const SellMatrix& A, const VT* p, VT* y, Workspace& w, int with_alpha, const SpmvTune& t, hipStream_t s);
// partials written by spmv_dot (with_alpha == 2: the partials only, no finalisation)
template <typename VT> int spmv_partials(const SellMatrix& A, const SpmvTune& t);
// r = b - A x in FP64, ||r||^2 -> out; optional rs = scale .* r (FP32)
// r = b - A x in FP64; dots ||r||^2 -> out[0], ||scale r||^2 -> out[1]; optional
// rs = scale .* r (FP32).
int spmv_residual64(const SellMatrix& A, const double* x, const double* b, double* r, float* rs, const double* scale,
                    Workspace& w, double* out, const SpmvTune& t, hipStream_t s);
// V-cycle kernels: VT is the stored value type (double, float or _Float16),
// T the vector/arithmetic type (double for VT = double, float otherwise).
template <typename VT, typename T>
int spmv_cheb(const SellMatrix& A, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, double c1, double c2,
              const SpmvTune& t, hipStream_t s);
template <typename VT, typename T>
int spmv_defect(const SellMatrix& A, const T* x, const T* b, T* res, const SpmvTune& t, hipStream_t s);
template <typename VT, typename T> int spmv_addto(const SellMatrix& A, const T* e, T* x, const SpmvTune& t,
                                                  hipStream_t s);
template <typename VT, typename T>
int spmv_restrict(const SellMatrix& A, const T* x, T* y, const SpmvTune& t, hipStream_t s);
// y = R x and the next level's first Chebyshev iterate x1 = d1 = c2 dinv .* y
template <typename VT, typename T>
int spmv_restrict_first(const SellMatrix& A, const T* x, T* y, const T* dinv, T* x1, T* d1, double c2,
                        const SpmvTune& t, hipStream_t s);

// standard block layout <-> sliced layout, and assembly
int sell_from_std(SellMatrix& A, const double* vals_std, int to64, int to32, hipStream_t s, int to16 = 0);
int sell_to_std(const SellMatrix& A, double* vals_std, hipStream_t s);
int assemble_gather(SellMatrix& A, const int* d_rowidx, const int* d_colidx, const long long* cptr, const int* codes,
                    const double* Ke, const unsigned char* node_fix, double* vals_std, hipStream_t s);
// Square element matrices Ke[M][npe*b][npe*b] (b = block size of A) into A's
// FP64 values through the contribution lists; sym != 0 assembles the
// symmetric part 0.5 (Ke + Ke^T); node_fix bit c marks DOF c of a node fixed.
int assemble_elem(SellMatrix& A, int npe, int sym, const int* d_rowidx, const int* d_colidx, const long long* cptr,
                  const int* codes, const double* Ke, const unsigned char* node_fix, double* vals_std, hipStream_t s);
int extract_diag(const SellMatrix& A, const long long* diag_block, double* diag, hipStream_t s);
int equilibrate(SellMatrix& A, const int* d_rowidx, const int* d_colidx, const double* scale, int want64e,
                hipStream_t s, int want16 = 0);

}  // namespace wn
