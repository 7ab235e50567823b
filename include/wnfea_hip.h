/*
 * libwnfea_hip.so -- MI355X (gfx950) solve path of WNFEA's C3D10 solver.
 *
 * Conventions
 *   - Return value: 0 ok, <0 argument/state error (-1 bad argument, -2 bad
 *     size, -3 bad mode, -4 not ready, -5 bad mesh), >0 a hipError_t code.
 *     No entry point exits or aborts.
 *   - "device" arrays are device pointers (e.g. torch tensors' data_ptr());
 *     "host" arrays are host pointers.  `stream` is a hipStream_t (NULL =
 *     default stream); kernels are enqueued on it.
 *   - Handles own their device buffers.  A handle is not thread-safe: use one
 *     handle per thread.
 *   - Block matrices use the standard block-CSR layout at the boundary
 *     (rowptr[nb+1], colidx[nnzb], values[nnzb][br*bc] row-major blocks) and
 *     a sliced-ELL layout inside the library.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository).  The reference's stateless exports copy host
 * arrays in and out per call; these take device arrays and a stream instead.
 */
#ifndef WNFEA_HIP_H
#define WNFEA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Library ABI version (1).  No reference counterpart. */
int wnfea_abi_version(void);

/* out[0] CUs, out[1] wavefront size, out[2] gfx arch number (0x950 = 2384),
 * out[3] L2 bytes, out[4] LDS bytes per block, out[5] HIP runtime version.
 * Replaces hip_get_device_summary (wnfea/solver/native/hip_kernels.cpp:1541). */
int wnfea_device_info(long long *out);

/* HBM streaming-read probe: a kernel of 512-thread blocks that reads the
 * buffer a (device, bytes a multiple of 16, 16-byte aligned) once with
 * 16-byte non-temporal loads, eight in flight per thread, and writes each
 * thread's sum of the floats it read to out[grid * 512] (device).  The rate
 * it reaches is the practical ceiling of the bandwidth-bound kernels, next
 * to the datasheet peak.  No reference counterpart (a measurement tool). */
int wnfea_hbm_stream_read(const float *a, long long bytes, float *out, int grid, void *stream);

/* ---------------------------------------------------------------- elements */

/* Batched C3D10 element stiffness, FP64.  conn[M][10] int32 (device),
 * nodes[N][3], E[M], nu[M] (device) -> Ke[M][30][30] (device, DOF 3*node+comp).
 * *bad (device int, zeroed by the caller) receives e+1 of an element with
 * det J <= 0.  Replaces the per-element element_stiffness_c3d10 loop of
 * assemble_c3d10_sparse_3dof (wnfea/elements/c3d10.py:219-249,
 * wnfea/solver/solid_solver.py:72-82). */
int wnfea_c3d10_ke(int M, const int *conn, const double *nodes, const double *E, const double *nu, double *Ke,
                   int *bad, void *stream);

/* Kernel behind wnfea_c3d10_ke for this process: 0 (default) the vector-ALU
 * contraction (one lane per upper 3x3 block pair), 1 the matrix-core
 * contraction Ke = B^T (w D B) with v_mfma_f64_16x16x4f64.  Both give Ke to
 * FP64 rounding and store it as contiguous 16-byte chunks.  No reference
 * counterpart (a kernel choice, as the reference's launch-configuration
 * helper get_optimal_launch_config, wnfea/solver/fast_kernels.py:243-260). */
int wnfea_c3d10_ke_variant(int variant);

/* Matrix-free internal forces F = sum_e K_e u_e, FP64, deterministic.
 * U and F have `stride` (3 or 6) doubles per node (device); fe_work[M][30]
 * is device scratch; n2e_ptr/n2e is the node->(element*10+local) incidence
 * from wnfea_node_incidence (device); accumulate != 0 adds to F.
 * Replaces hip_compute_c3d10_forces (wnfea/solver/native/hip_kernels.cpp:1633)
 * and compute_c3d10_internal_forces_native (wnfea/solver/native/kernels.cpp:244). */
int wnfea_c3d10_fint(int M, int n_nodes, const int *conn, const double *nodes, const double *E, const double *nu,
                     const double *U, int stride, const int *n2e_ptr, const int *n2e, double *fe_work, double *F,
                     int accumulate, void *stream);

/* Matrix-free C3D10 operator K (FP64) with the assembled matrix's Dirichlet
 * rule (a DOF flagged in node_fix, bit c of its node's byte, keeps a unit
 * diagonal and no coupling), stored as per-Gauss-point inverse Jacobians
 * and weights (320 B per element) instead of K (about 3 KB per element of
 * FP64 blocks): the reference's matrix-free f_int kernel used as the
 * operator of a solve (wnfea/solver/native/hip_kernels.cpp:283-425,
 * c3d10_internal_forces_kernel).  The mesh arrays (device: conn, nodes, E,
 * nu, node_fix (may be null), n2e_ptr/n2e from wnfea_node_incidence) stay
 * owned by the caller and must outlive the handle.  *bad = e + 1 for an
 * element with det J <= 0 (left 0 otherwise; device int). */
int wnfea_mf_create(int M, int n_nodes, const int *conn, const double *nodes, const double *E, const double *nu,
                    const unsigned char *node_fix, const int *n2e_ptr, const int *n2e, int *bad, void **handle,
                    void *stream);
void wnfea_mf_destroy(void *handle);

/* 1 when every element is affine (straight edges, midside nodes at the edge
 * midpoints: J is constant, so one inverse Jacobian per element is kept, 80
 * B instead of 320 B), 0 for a mesh with curved elements, < 0 on a bad
 * handle.  No reference counterpart. */
int wnfea_mf_affine(void *handle);

/* Patch layout of the operator's products (affine meshes; built by
 * wnfea_mf_create with 64 elements per patch when it fits): elements in
 * Morton order of their centroids, cut into patches of `elems` (a multiple
 * of 64, at most 256); one workgroup per patch stages the patch's nodal
 * values and element forces in LDS and writes one partial sum per (patch,
 * node), and a node-ordered kernel adds each node's partials in patch order.
 * elems == 0 drops the layout (per-element contributions and a node gather,
 * as the reference's c3d10_internal_forces_kernel + gather,
 * wnfea/solver/native/hip_kernels.cpp:283-425).  -3 for a curved mesh, -2
 * when an FP64 patch would not fit a workgroup's LDS.  Hierarchies attached
 * before use the per-element path until attached again.  No reference
 * counterpart. */
int wnfea_mf_set_patch(void *handle, int elems, void *stream);

/* out[0..5] (host) = elements per patch (0: no patch layout), patches, most
 * local nodes of one patch, local nodes of all patches, affine (1/0), the
 * per-patch stride of the local-node arrays.  No reference counterpart. */
int wnfea_mf_info(void *handle, long long *out);

/* y = K u, device FP64 vectors of 3 * n_nodes (matrix-free, deterministic). */
int wnfea_mf_apply(void *handle, const double *u, double *y, void *stream);

/* r = f - K u (matrix-free) and out[0] = ||r||, out[1] = ||f|| (host). */
int wnfea_mf_residual(void *solver, void *handle, const double *u, const double *f, double *r, double *out,
                      void *stream);

/* FP64 PCG on the matrix-free operator from x = 0; pkind, dinv, amg, amg_vt
 * and scale as in wnfea_pcg (the AMG hierarchy still comes from an assembled
 * operator).  Replaces the operator of the reference's solves,
 * wnfea/solver/solid_solver.py:197-308. */
int wnfea_pcg_mf(void *solver, void *handle, int pkind, const double *dinv, void *amg, int amg_vt,
                 const double *scale, const double *b, double *x, double rtol, int max_iter, int check_every,
                 int *iters, double *rel, void *stream);

/* Make the finest level's smoothing steps and defect of the AMG V-cycles
 * with FP64 or FP32 finest values (amg_vt 0, 1, 4, 5) matrix-free: each
 * product is diag(scale) K diag(scale) x with K this matrix-free operator
 * (FP32 cycles use a rounded copy of its quadrature data) and scale the
 * device FP64 equilibration of the fine matrix (3 * n_nodes), followed by the
 * step's own update in node order.  Replaces the finest-level smoother and
 * defect products of the reference's V-cycle (wnfea/solver/native/
 * hip_kernels.cpp:525-649 jacobi_smooth_csr_*, used at 2106-2160 and
 * 2219-2240).  handle = NULL detaches; the operator and scale must outlive
 * the attachment. */
int wnfea_amg_attach_mf(void *amg, void *handle, const double *scale, void *stream);

/* Gauss-point stresses (optional gp_sig[M][4][6], gp_vm[M][4]), cell means
 * cell_sig[M][6], cell_vm[M], and (if nodal_vm) the nodal average of cell_vm.
 * U has 3 doubles per node.  Replaces element_stresses_c3d10 and
 * compute_c3d10_stress_field (wnfea/elements/c3d10.py:286-318,
 * wnfea/solver/solid_solver.py:311-365). */
int wnfea_c3d10_stress(int M, int n_nodes, const int *conn, const double *nodes, const double *E, const double *nu,
                       const double *U, const int *n2e_ptr, const int *n2e, double *gp_sig, double *gp_vm,
                       double *cell_sig, double *cell_vm, double *nodal_vm, int *bad, void *stream);

/* Adjoint sensitivity with respect to per-element Young's modulus:
 * grad[e] = -(1/E_e) lam_e^T K_e u_e (K is linear in E_e).  u, lam 3 per node.
 * No reference counterpart (the reference has no sensitivities); it serves the
 * compliance / tip-displacement gradients of the adjoint configuration. */
int wnfea_adjoint_grad(int M, const int *conn, const double *nodes, const double *E, const double *nu,
                       const double *u, const double *lam, double *fe_work, double *grad, void *stream);

/* ---------------------------------------------------------------- co-rotational beams */

/* Co-rotational internal forces of 2-node beams, FP64, deterministic.
 * conn[M][2], nodes[N][3], props[M][6] = (E, G, A, Iy, Iz, J), U[N][6]
 * (device); fe_work[M][12] device scratch; n2e_ptr/n2e the node incidence
 * (wnfea_node_incidence with npe = 2); F[N][6] (accumulate != 0 adds).
 * Replaces hip_compute_beam_forces (wnfea/solver/native/hip_kernels.cpp:53-241),
 * compute_corotational_element_forces (wnfea/solver/corotational_beam.py:178-233)
 * and the vectorised beam path of compute_internal_forces_fast
 * (wnfea/solver/fast_kernels.py:451-537). */
int wnfea_beam_forces(int M, int n_nodes, const int *conn, const double *nodes, const double *props, const double *U,
                      const int *n2e_ptr, const int *n2e, double *fe_work, double *F, int accumulate, void *stream);

/* Element tangents Kt[M][12][12] = d f_e / d u_e of the co-rotational
 * forces: h_rel <= 0 gives the exact (consistent) derivative, by forward-mode
 * differentiation of the force routine in dual numbers; h_rel > 0 central
 * differences (h = h_rel * L0 for translations, h_rel for rotations).
 * Replaces the finite-difference Jacobian action of
 * MatrixFreeJFNKOperator._matvec (wnfea/solver/jfnk_operator.py:98-134) with
 * an assembled tangent. */
int wnfea_beam_tangent(int M, const int *conn, const double *nodes, const double *props, const double *U,
                       double h_rel, double *Kt, void *stream);

/* Linear 12x12 beam stiffness T^T K_local T per element (device).  Replaces
 * build_element_stiffness_3d_beam and build_transformation_matrix
 * (wnfea/solver/assembler.py:13-148), frame rules included. */
int wnfea_beam_ke_linear(int M, const int *conn, const double *nodes, const double *props, double *Ke, void *stream);

/* ---------------------------------------------------------------- host setup */

/* Reverse Cuthill-McKee order (host arrays): perm[new] = old, for a symmetric
 * CSR graph with sorted columns and the seed order numpy.argsort(degrees)
 * uses.  Replaces scipy.sparse.csgraph.reverse_cuthill_mckee as called by
 * compute_rcm_order (wnfea/mesh/rcm.py:73-92). */
int wnfea_rcm(int n, const int *ptr, const int *ind, const int *seed_order, int *perm);

/* Node graph of an element mesh (host arrays; npe nodes per element): sorted
 * unique neighbours including the node itself.  With node_free (host uint8),
 * off-diagonal pairs touching a fully fixed node are dropped (the Dirichlet
 * pattern rule).  Returns nnz, or -1 if cap is too small.  Replaces
 * compute_node_adjacency_matrix (wnfea/mesh/rcm.py:20-70) and the COO->CSR
 * pattern + eliminate_zeros of assemble_c3d10_sparse_3dof
 * (wnfea/solver/solid_solver.py:50-64,84,95-127). */
long long wnfea_node_graph(int n_nodes, int M, int npe, const int *conn, const unsigned char *node_free, int *rowptr,
                           int *colidx, long long cap);

/* Per-block contribution lists for deterministic assembly (host arrays):
 * cptr[nnzb+1] (int64), codes = e*npe*npe + a*npe + b in element order.
 * Returns the number of codes, or -1 if cap is too small.  Replaces the COO
 * duplicate summation of assemble_c3d10_sparse_3dof
 * (wnfea/solver/solid_solver.py:62-84). */
long long wnfea_contributions(int n_nodes, int M, int npe, const int *conn, const int *rowptr, const int *colidx,
                              long long *cptr, int *codes, long long cap);

/* Node -> (element*npe + local) incidence in element order (host arrays).
 * Serves the deterministic gathers of wnfea_c3d10_fint/_stress; replaces the
 * atomic scatter of c3d10_internal_forces_kernel
 * (wnfea/solver/native/hip_kernels.cpp:418-423). */
int wnfea_node_incidence(int n_nodes, int M, int npe, const int *conn, int *ptr, int *codes);

/* Smoothed-aggregation aggregates of a strength graph (host; CSR without the
 * diagonal).  agg[i] = aggregate id or -1 (isolated).  Returns the count.
 * Replaces aggregate_beam_nodes (wnfea/solver/amg_preconditioner.py:72-124). */
long long wnfea_sa_aggregate(int n, const int *ptr, const int *ind, int *agg);

/* Tentative prolongator (host): per-aggregate orthonormalisation of the
 * near-null-space B[n_nodes*b][k]; P as block CSR (b x k blocks, one per
 * aggregated node) and the coarse near-null-space Bc[n_agg*k][k].  Replaces
 * the per-aggregate QR of _build_hierarchy
 * (wnfea/solver/amg_preconditioner.py:241-281). */
int wnfea_sa_tentative(int n_nodes, int b, int k, const int *agg, int n_agg, const double *B, int *p_rowptr,
                       int *p_colidx, double *p_vals, double *Bc);

/* Host block-sparse kernels of the smoothed-aggregation set-up (host
 * pointers, host threads; `threads` 0 = OMP_NUM_THREADS when set, else one
 * per hardware thread).  They
 * replace the scipy sparse products, transposes and matrix-vector products
 * of the reference's setup (wnfea/solver/amg_preconditioner.py:241-293,
 * its Galerkin product R A P at :283-293).
 *
 * C = A B (block CSR, row-major blocks ra x ca and ca x cb, nc block
 * columns of B).  With c_idx == NULL: structure pass, c_ptr[0..nr] gets the
 * prefix sums of C's row block counts and the total is returned; then call
 * again with c_idx / c_val of that size.  Columns come out sorted; every
 * block is summed in a fixed order (independent of the thread count). */
long long wnfea_host_bsr_spgemm(int nr, int ra, int ca, int cb, int nc, const int *a_ptr, const int *a_idx,
                                const double *a_val, const int *b_ptr, const int *b_idx, const double *b_val,
                                int *c_ptr, int *c_idx, double *c_val, int threads);

/* T = A^T (t_ptr[nc + 1], t_idx / t_val[nnzb], blocks transposed). */
int wnfea_host_bsr_transpose(int nr, int nc, int ra, int ca, const int *a_ptr, const int *a_idx,
                             const double *a_val, int *t_ptr, int *t_idx, double *t_val);

/* In place A = (A + A^T) / 2 for a square block CSR with a structurally
 * symmetric pattern and sorted columns; zero diagonal entries become 1 (the
 * reference's fix of vanished coarse DOFs, amg_preconditioner.py:288-293).
 * Returns the number of diagonal entries set, or -1 if a mirror block is
 * missing. */
long long wnfea_host_bsr_symmetrize(int n, int b, const int *ptr, const int *idx, double *val, int threads);

/* y = A x for a square block CSR (b <= 8) on host threads. */
int wnfea_host_bsr_spmv(int n, int b, const int *ptr, const int *idx, const double *val, const double *x, double *y,
                        int threads);

/* Host set-up of the matrix-free patch layout (host pointers; what
 * wnfea_mf_set_patch builds and uploads).  order[M] = the elements grouped
 * by connected component (elements sharing a corner; components in the
 * order of their first element) and inside a component sorted by the Morton
 * code of their corner centroids (21 bits per axis on the component's
 * bounding cube; ties in element order).  No reference counterpart (the
 * reference's matrix-free kernel walks elements in mesh order,
 * wnfea/solver/native/hip_kernels.cpp:283-425). */
int wnfea_host_morton_order(int M, const int *conn, const double *nodes, int *order);

/* Patches of `elems` consecutive slots of `order` (slot s holds element
 * order[s]; a patch's local nodes are its elements' distinct nodes,
 * ascending).  wnfea_host_mf_patch_sizes returns the number of local nodes
 * of all patches and *lmax, the most of one patch.  wnfea_host_mf_patches
 * fills, for stride > lmax: pnode[p stride + j] the global node of local
 * node j of patch p (-1 past the last); pconn[10 M] the local node of corner
 * i of slot s; pslot[10 M] per patch the local nodes' slot codes
 * 10 (s - p elems) + i, each list ascending, local node j's list from
 * poff[p stride + j] to poff[p stride + j + 1] (entries past the last node
 * hold the end); nhead[4 n_nodes] node n's partial slots p stride + j in
 * patch order, the first three (-1 where absent) and their count, or with
 * more than three -(k + 1), where novf[k] is the count of the rest and
 * novf[k + 1 ...] the rest (novf needs room for the local-node count + 1).
 * Returns the length of novf. */
long long wnfea_host_mf_patch_sizes(int M, int n_nodes, const int *conn, const int *order, int elems, int *lmax);
long long wnfea_host_mf_patches(int M, int n_nodes, const int *conn, const int *order, int elems, int stride,
                                int *pnode, unsigned short *pconn, unsigned short *poff, unsigned short *pslot,
                                int *nhead, int *novf);

/* ---------------------------------------------------------------- block matrices */

/* Create a br x bc block matrix from its host block-CSR structure; sigma is
 * the row-sorting window of the sliced-ELL layout (multiple of 64; <=0 =
 * 4096); lanes is the number of wavefront lanes per block row (power of two
 * <= 64; <=0 = chosen from the shape).  Returns NULL on failure.  Replaces BSR6x6Matrix / csr_to_bsr6x6
 * (wnfea/solver/bsr_matrix.py:19-122) and the device uploads of
 * hip_spmv_bsr6x6_fp32/fp64 (wnfea/solver/native/hip_kernels.cpp:2975-3087). */
void *wnfea_bsr_create(int br, int bc, int n_rows, int n_cols, const int *rowptr, const int *colidx, int sigma,
                       int lanes);
void wnfea_bsr_destroy(void *h);

/* out[0..8]: n_slices, slots, nnzb, sigma, has FP64, has FP32, has FP64 equilibrated, lanes per row,
 * has FP16. */
int wnfea_bsr_info(void *h, long long *out);

/* SpMV launch shape: waves per slice (1, 2, 4), XCD-contiguous slice mapping,
 * blocks in flight per lane (1, 2, 4).  No reference counterpart (the
 * reference's launch sizes are fixed heuristics, wnfea/solver/fast_kernels.py:243-260). */
int wnfea_bsr_tune(void *h, int waves, int xcd_remap, int unroll);

/* The same for one value precision only (prec 0: FP64 values, 1: FP32 and
 * FP16 values): the fastest shapes differ between the precisions.  No
 * reference counterpart (wnfea/solver/fast_kernels.py:243-260 has one fixed
 * launch size). */
int wnfea_bsr_tune_precision(void *h, int prec, int waves, int xcd_remap, int unroll);

/* Values from the standard layout (device FP64 [nnzb][br*bc]); which: 1 FP64, 2 FP32, 3 both. */
int wnfea_bsr_set_values(void *h, const double *vals_std, int which, void *stream);

/* FP64 values back in the standard layout (device).  Replaces bsr6x6_to_csr's
 * value source (wnfea/solver/bsr_matrix.py:125-152). */
int wnfea_bsr_get_values(void *h, double *vals_std, void *stream);

/* Assemble 3x3 blocks from Ke (device) via the contribution lists (device
 * rowidx/colidx per block, cptr, codes), applying the Dirichlet rule per
 * scalar entry from node_fix (device uint8 bits per node; NULL = no BCs);
 * optionally also writes vals_std.  Replaces assemble_c3d10_sparse_3dof's
 * value path (wnfea/solver/solid_solver.py:66-127). */
int wnfea_bsr_assemble(void *h, const int *rowidx, const int *colidx, const long long *cptr, const int *codes,
                       const double *Ke, const unsigned char *node_fix, double *vals_std, void *stream);

/* Element matrices Ke[M][npe*b][npe*b] of b x b-block matrix h (b = 3 or 6)
 * through the contribution lists (codes e*npe*npe + a*npe + c); sym != 0
 * assembles 0.5 (Ke + Ke^T); node_fix bits as wnfea_bsr_assemble (bit c =
 * DOF c fixed).  Replaces the beam branch of assemble_global_system_sparse
 * (wnfea/solver/assembler.py:285-450) and hip_assemble_beam_system_csr
 * (wnfea/solver/fast_kernels.py:666-763). */
int wnfea_bsr_assemble_elem(void *h, int npe, int sym, const int *rowidx, const int *colidx, const long long *cptr,
                            const int *codes, const double *Ke, const unsigned char *node_fix, double *vals_std,
                            void *stream);

/* Scalar diagonal (device) given diag_block[i] = standard index of block (i,i). */
int wnfea_bsr_diag(void *h, const long long *diag_block, double *diag, void *stream);

/* FP32 values of S K S with S = diag(scale) (device), plus the FP64 copy
 * (want64e) and an FP16 copy (want16) for lower-precision preconditioning.
 * Replaces the equilibration of BlockBeamAMGPreconditioner
 * (wnfea/solver/amg_preconditioner.py:180-185) and its FP16 value copies
 * (wnfea/solver/native/hip_kernels.cpp:1924-2060, precision_mode 2). */
int wnfea_bsr_equilibrate(void *h, const int *rowidx, const int *colidx, const double *scale, int want64e,
                          int want16, void *stream);

/* y = alpha A x + beta y; prec 0 FP64, 1 FP32 (FP32 vectors), 2 FP64
 * equilibrated.  Replaces hip_spmv_bsr6x6_fp32 / _fp64
 * (wnfea/solver/native/hip_kernels.cpp:2975,3032) and bsr6x6_spmv_cpu
 * (wnfea/solver/bsr_matrix.py:155-193). */
int wnfea_bsr_spmv(void *h, int prec, const void *x, void *y, double alpha, double beta, void *stream);

/* y = A x and *out = x . (A x) (FP64 reduction, host scalar; synchronises).
 * Replaces spmv_csr_dot_kernel's fused product (hip_kernels.cpp:1171-1228). */
int wnfea_bsr_spmv_dot(void *h, int prec, const void *x, void *y, double *out, void *stream);

/* ---------------------------------------------------------------- AMG */

/* Replaces hip_amg_create / hip_amg_set_level / hip_amg_set_coarse_cholesky /
 * hip_amg_apply_vcycle / hip_amg_destroy
 * (wnfea/solver/native/hip_kernels.cpp:1891,1924,2061,2519,2541). */
/* store: value precisions kept on every level, bit0 FP64, bit1 FP32, bit2
 * FP16 (the reference's precision_mode 0/1/2). */
void *wnfea_amg_create(int n_levels, int store);
void wnfea_amg_destroy(void *h);

/* Level data (host arrays).  Level 0 passes fine_bsr (its FP32 / FP64
 * equilibrated values are the operator) and NULL A arrays; other levels pass
 * A as br x br block CSR.  dinv: point-Jacobi inverse diagonal; Chebyshev on
 * [lmax/ratio, lmax] of the given degree; P (br x bc_coarse blocks) and
 * R = P^T (bc_coarse x br) except on the coarsest level. */
int wnfea_amg_set_level(void *h, int lvl, void *fine_bsr, int br, int n_rows, const int *a_rowptr,
                        const int *a_colidx, const double *a_vals, const double *dinv, double lmax, double ratio,
                        int degree, int bc_coarse, int n_coarse_rows, const int *p_rowptr, const int *p_colidx,
                        const double *p_vals, const int *r_rowptr, const int *r_colidx, const double *r_vals,
                        int sigma);

/* Coarsest-level dense inverse (host, n x n row-major FP64). */
int wnfea_amg_set_coarse(void *h, int n, const double *inv);

/* Allocate the cycle vectors of the stored precisions and set the SpMV
 * launch shape (as wnfea_bsr_tune). */
int wnfea_amg_finalize(void *h, int waves, int xcd_remap, int unroll);

/* SpMV launch shape of the cycle for one value precision (prec as in
 * wnfea_bsr_tune_precision), after wnfea_amg_finalize.  No reference
 * counterpart (wnfea/solver/fast_kernels.py:243-260). */
int wnfea_amg_tune_precision(void *h, int prec, int waves, int xcd_remap, int unroll);

/* z = M r, one V-cycle (device vectors).  vt: 0 FP64 values and vectors;
 * 1 FP32 values, 2 FP16 values on every level, 3 FP16 values on the finest
 * level and FP32 below, 4 FP32 on the finest level and FP16 below, 5 FP16
 * values in the smoothing steps and FP32 in the defects and transfers (not
 * symmetric; wnfea_pcg then uses the flexible beta) (vt >= 1: FP32 vectors
 * and arithmetic). */
int wnfea_amg_vcycle(void *h, int vt, const void *r, void *z, void *stream);

/* Coarsest-level solve of the FP32 / FP16 cycles: fp64 != 0 applies the
 * FP64 dense inverse (FP64 products, FP32 vectors in and out; needs FP64
 * storage), 0 (default) its FP32 copy.  No reference counterpart (the
 * reference factorises its coarse level in the cycle's working precision,
 * wnfea/solver/amg_preconditioner.py:304-330). */
int wnfea_amg_set_coarse_fp64(void *h, int fp64);

/* reps Chebyshev smoothing steps of level lvl (0 <= lvl < last) on the
 * level's own cycle vectors, with its assembled values in FP64 (vt 0) or
 * FP32 (vt 1): the smoother kernel of the V-cycle alone, so that its time per
 * launch and its HBM rate can be measured.  No reference counterpart (the
 * reference times only its SpMV, hip_spmv_bsr6x6_benchmark,
 * wnfea/solver/native/hip_kernels.cpp:3089-3164). */
int wnfea_amg_level_cheb(void *h, int lvl, int vt, int reps, void *stream);

/* ---------------------------------------------------------------- solvers */

/* Solver workspace (Krylov vectors, reduction buffers, the replayable
 * iteration graph).  Replaces the per-handle scratch of hip_amg_solve_pcg
 * (wnfea/solver/native/hip_kernels.cpp:2384-2517). */
void *wnfea_solver_create(void);
void wnfea_solver_destroy(void *h);

/* enable != 0: full chunks of check_every PCG iterations are captured once
 * as a HIP graph and replayed (bitwise the same kernels as host launches);
 * 0 (default): every kernel is launched from the host.  No reference
 * counterpart (its loop launches every kernel). */
int wnfea_solver_set_graphs(void *h, int enable);

/* PCG from x = 0 until ||r|| <= rtol ||b|| (recursive residual); prec as
 * wnfea_bsr_spmv; pkind 0 none, 1 Jacobi (dinv in the vectors' precision),
 * 2 AMG (amg_vt as wnfea_amg_vcycle; with scale = S the preconditioner is
 * z = S M(S r), M the V-cycle of S K S, so an FP64 loop on K takes an FP32 or
 * FP16 cycle).  Iterations run in device-scalar chunks of check_every between
 * host checks.  Replaces hip_amg_solve_pcg(_device)
 * (wnfea/solver/native/hip_kernels.cpp:2384-2517) and solve_c3d10_pcg's FP64
 * branch (wnfea/solver/solid_solver.py:302-308). */
int wnfea_pcg(void *solver, void *bsr, int prec, int pkind, const void *dinv, void *amg, int amg_vt,
              const double *scale, const void *b, void *x, double rtol, int max_iter, int check_every, int *iters,
              double *rel, void *stream);

/* FP64 iterative refinement: r = f - K u (FP64), inner FP32 PCG on S K S to
 * inner_rtol, u += S d, until ||f - K u|| <= tol ||f||.  Replaces
 * solve_c3d10_pcg's mixed_fp32 branch (wnfea/solver/solid_solver.py:271-300). */
int wnfea_ir(void *solver, void *bsr, const double *scale, int pkind, const float *dinv32, void *amg, int amg_vt,
             const double *f, double *u, double tol, double inner_rtol, int max_outer, int max_inner,
             int check_every, int zero_init, int *outer, int *inner, double *rel, void *stream);

/* Multiple right-hand sides: PCG on nrhs (4 or 8) columns at once, B and X
 * FP64 [n][nrhs] interleaved (device).  Per-column CG (rho, alpha, beta,
 * ||r||^2 per column; converged columns are frozen), one FP64 SpMM per
 * iteration (K read once for all columns), the FP32 V-cycle (amg_vt 1) on
 * S K S for all columns, z = S M(S r).  iters[nrhs], rel[nrhs] (host) get
 * the per-column iterations and recursive relative residuals.  No reference
 * counterpart (the reference solves one load case per call,
 * wnfea/solver/solid_solver.py:197-308). */
int wnfea_pcg_multi(void *solver, void *bsr, void *amg, int amg_vt, const double *scale, int nrhs, const double *B,
                    double *X, double rtol, int max_iter, int check_every, int *iters, double *rel, void *stream);

/* Y = K X on nrhs (4 or 8) interleaved columns; prec 0 FP64, 1 FP32 values.
 * Multi-vector form of hip_spmv_bsr6x6_fp32/_fp64
 * (wnfea/solver/native/hip_kernels.cpp:2975,3032). */
int wnfea_bsr_spmm(void *h, int prec, int nrhs, const void *X, void *Y, void *stream);

/* r = f - K u (FP64; r may be NULL) and out[0] = ||r||, out[1] = ||f||. */
int wnfea_residual(void *solver, void *bsr, const double *u, const double *f, double *r, double *out, void *stream);

/* wnfea_residual without any allocation, host copy or synchronization:
 * r = f - K u (FP64; r may be NULL) and out_dev[0] = ||r||, out_dev[1] = ||f||
 * written to DEVICE memory.  The reduction partials live in a static
 * per-GPU scratch array of the library (2 MiB, held by each GPU it runs on
 * until the process ends), so a matrix of more than 65 528 sliced-ELL slices
 * (about 1M block rows at 16 rows per slice) is rejected with -2 (bad size).
 * Calls that overlap on one GPU share that scratch and race.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops). */
int wnfea_residual_noalloc(void *bsr, const double *u, const double *f, double *r, double *out_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* WNFEA_HIP_H */
