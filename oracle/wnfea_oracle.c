/*
 * CPU FP64 oracle for the C3D10 solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C restatement of the reference algorithms, used as the
 * checker of the HIP library.  Only the tests, the smoke check and the bench's
 * CPU-baseline leg load it; the product library never links or calls it.
 *
 * Function                    follows (reference file:line)
 * ------------------------------------------------------------------------------
 * or_shape_derivs             wnfea/elements/c3d10.py:35-106 (dN/dxi, Abaqus order)
 * or_c3d10_ke                 wnfea/elements/c3d10.py:140-249 (B 6x30, C, sum w B^T C B detJ, symmetrise)
 * or_c3d10_fint               wnfea/elements/c3d10.py:252-283
 * or_c3d10_stress             wnfea/elements/c3d10.py:286-318
 * or_fint_global_stride       wnfea/solver/native/kernels.cpp:244-376 (U/F stride 6, skip |detJ|<1e-15)
 * or_ke_batch                 per-element loop of wnfea/solver/solid_solver.py:72-82
 * or_csr_accumulate           coo->csr duplicate summation of solid_solver.py:84 (element order)
 * or_csr_spmv                 y = K x of scipy CSR
 * or_pcg                      scipy cg with Jacobi M (solid_solver.py:232-236, 302-308)
 * or_ir_mixed                 mixed_fp32 refinement loop, solid_solver.py:271-300
 * or_rcm                      scipy.sparse.csgraph.reverse_cuthill_mckee called at
 *                             wnfea/mesh/rcm.py:85 (scipy 1.15.3 algorithm; seed order
 *                             supplied by the caller = numpy argsort of the degrees)
 * or_stress_field             wnfea/solver/solid_solver.py:311-365
 *
 * Pinning: each function is checked in tests/ against the reference's own
 * known answers (tests of test_c3d10.py: delta property, partition of unity,
 * 6 zero / 24 positive eigenvalues, patch test, Ke u = f_int), against
 * oracle/_ref (the reference's kernels.cpp compiled from source) for the
 * stride-6 internal forces, against scipy (spsolve, reverse_cuthill_mckee,
 * CSR products) and against closed forms (PL^3/3EI-type limits are not used as
 * pins, only as sanity prints).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_EXPORT __attribute__((visibility("default")))

static const double OR_W = 1.0 / 24.0;

static void gauss_points(double gp[4][3]) {
    const double a = (5.0 - sqrt(5.0)) / 20.0;
    const double b = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
    const double pts[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    memcpy(gp, pts, sizeof(pts));
}

/* dN/dxi (3x10) at (xi, eta, zeta); c3d10.py:35-106 */
OR_EXPORT void or_shape_derivs(double xi, double eta, double zeta, double N[10], double dN[3][10]) {
    const double L1 = 1.0 - xi - eta - zeta, L2 = xi, L3 = eta, L4 = zeta;
    const double n[10] = {L1 * (2 * L1 - 1), L2 * (2 * L2 - 1), L3 * (2 * L3 - 1), L4 * (2 * L4 - 1),
                          4 * L1 * L2, 4 * L2 * L3, 4 * L3 * L1, 4 * L1 * L4, 4 * L2 * L4, 4 * L3 * L4};
    /* dN_dL[k][i] = dN_i / dL_k */
    double dL[4][10];
    memset(dL, 0, sizeof(dL));
    dL[0][0] = 4 * L1 - 1; dL[1][1] = 4 * L2 - 1; dL[2][2] = 4 * L3 - 1; dL[3][3] = 4 * L4 - 1;
    dL[0][4] = 4 * L2; dL[1][4] = 4 * L1;
    dL[1][5] = 4 * L3; dL[2][5] = 4 * L2;
    dL[0][6] = 4 * L3; dL[2][6] = 4 * L1;
    dL[0][7] = 4 * L4; dL[3][7] = 4 * L1;
    dL[1][8] = 4 * L4; dL[3][8] = 4 * L2;
    dL[2][9] = 4 * L4; dL[3][9] = 4 * L3;
    for (int i = 0; i < 10; ++i) {
        if (N) N[i] = n[i];
        dN[0][i] = -dL[0][i] + dL[1][i];
        dN[1][i] = -dL[0][i] + dL[2][i];
        dN[2][i] = -dL[0][i] + dL[3][i];
    }
}

static void elasticity(double E, double nu, double C[6][6]) {
    const double f = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double lam = f * nu, d = f * (1.0 - nu), G = E / (2.0 * (1.0 + nu));
    memset(C, 0, 36 * sizeof(double));
    C[0][0] = C[1][1] = C[2][2] = d;
    C[0][1] = C[1][0] = C[0][2] = C[2][0] = C[1][2] = C[2][1] = lam;
    C[3][3] = C[4][4] = C[5][5] = G;
}

/* B (6x30) and det J at a Gauss point (c3d10.py:109-189).  B is filled whenever det J != 0;
 * the return value is 0 for det J > 0 and -1 otherwise (the Python element raises then). */
static int bmatrix(const double X[10][3], const double gp[3], double B[6][30], double *detJ) {
    double dN[3][10];
    or_shape_derivs(gp[0], gp[1], gp[2], NULL, dN);
    double J[3][3] = {{0}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < 10; ++i) J[r][c] += dN[r][i] * X[i][c];
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    *detJ = det;
    if (det == 0.0 || det != det) return -1;
    const double id = 1.0 / det;
    const double inv[3][3] = {
        {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
        {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
        {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
    memset(B, 0, 6 * 30 * sizeof(double));
    for (int i = 0; i < 10; ++i) {
        double g[3];
        for (int r = 0; r < 3; ++r) g[r] = inv[r][0] * dN[0][i] + inv[r][1] * dN[1][i] + inv[r][2] * dN[2][i];
        const int c = 3 * i;
        B[0][c] = g[0];
        B[1][c + 1] = g[1];
        B[2][c + 2] = g[2];
        B[3][c] = g[1]; B[3][c + 1] = g[0];
        B[4][c + 1] = g[2]; B[4][c + 2] = g[1];
        B[5][c] = g[2]; B[5][c + 2] = g[0];
    }
    return det > 0.0 ? 0 : -1;
}

/* Ke (30x30 row-major, DOF 3*node+comp); c3d10.py:219-249. Returns -1 on det J <= 0. */
OR_EXPORT int or_c3d10_ke(const double *Xf, double E, double nu, double *Ke) {
    const double (*X)[3] = (const double (*)[3])Xf;
    double C[6][6], gp[4][3], B[6][30], CB[6][30];
    elasticity(E, nu, C);
    gauss_points(gp);
    memset(Ke, 0, 900 * sizeof(double));
    for (int g = 0; g < 4; ++g) {
        double detJ;
        if (bmatrix(X, gp[g], B, &detJ)) return -1;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 30; ++c) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += C[r][k] * B[k][c];
                CB[r][c] = s;
            }
        for (int i = 0; i < 30; ++i)
            for (int j = 0; j < 30; ++j) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += B[k][i] * CB[k][j];
                Ke[i * 30 + j] += OR_W * s * detJ;
            }
    }
    for (int i = 0; i < 30; ++i)
        for (int j = i + 1; j < 30; ++j) {
            const double s = 0.5 * (Ke[i * 30 + j] + Ke[j * 30 + i]);
            Ke[i * 30 + j] = Ke[j * 30 + i] = s;
        }
    return 0;
}

/* f_int = sum w B^T C B u detJ; c3d10.py:252-283 */
OR_EXPORT int or_c3d10_fint(const double *Xf, const double *ue, double E, double nu, double *f) {
    const double (*X)[3] = (const double (*)[3])Xf;
    double C[6][6], gp[4][3], B[6][30];
    elasticity(E, nu, C);
    gauss_points(gp);
    memset(f, 0, 30 * sizeof(double));
    for (int g = 0; g < 4; ++g) {
        double detJ, eps[6] = {0}, sig[6] = {0};
        if (bmatrix(X, gp[g], B, &detJ)) return -1;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 30; ++c) eps[r] += B[r][c] * ue[c];
        for (int r = 0; r < 6; ++r)
            for (int k = 0; k < 6; ++k) sig[r] += C[r][k] * eps[k];
        for (int c = 0; c < 30; ++c) {
            double s = 0.0;
            for (int r = 0; r < 6; ++r) s += B[r][c] * sig[r];
            f[c] += OR_W * s * detJ;
        }
    }
    return 0;
}

/* Gauss-point Cauchy stress (4x6) and von Mises (4); c3d10.py:286-318 */
OR_EXPORT int or_c3d10_stress(const double *Xf, const double *ue, double E, double nu, double *sig_out, double *vm_out) {
    const double (*X)[3] = (const double (*)[3])Xf;
    double C[6][6], gp[4][3], B[6][30];
    elasticity(E, nu, C);
    gauss_points(gp);
    for (int g = 0; g < 4; ++g) {
        double detJ, eps[6] = {0}, s[6] = {0};
        if (bmatrix(X, gp[g], B, &detJ)) return -1;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 30; ++c) eps[r] += B[r][c] * ue[c];
        for (int r = 0; r < 6; ++r)
            for (int k = 0; k < 6; ++k) s[r] += C[r][k] * eps[k];
        for (int r = 0; r < 6; ++r) sig_out[g * 6 + r] = s[r];
        vm_out[g] = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                                (s[2] - s[0]) * (s[2] - s[0]) + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
    }
    return 0;
}

static void gather_elem(const int *conn, const double *nodes, int e, double X[30]) {
    for (int i = 0; i < 10; ++i) {
        const int n = conn[e * 10 + i];
        X[3 * i] = nodes[3 * n]; X[3 * i + 1] = nodes[3 * n + 1]; X[3 * i + 2] = nodes[3 * n + 2];
    }
}

/* All element stiffnesses of a mesh: Ke_out[M][900]. Returns the first bad element + 1, or 0. */
OR_EXPORT int or_ke_batch(int M, const int *conn, const double *nodes, const double *E, const double *nu, double *Ke_out) {
    int bad = 0;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < M; ++e) {
        double X[30];
        gather_elem(conn, nodes, e, X);
        if (or_c3d10_ke(X, E[e], nu[e], Ke_out + (size_t)e * 900)) {
#pragma omp critical
            if (!bad || e + 1 < bad) bad = e + 1;
        }
    }
    return bad;
}

/* Global internal forces with arbitrary DOF stride (3 or 6); kernels.cpp:244-376 semantics
 * (skip Gauss points with |detJ| < 1e-15, accumulate into F). Serial => deterministic. */
OR_EXPORT void or_fint_global_stride(int M, const int *conn, const double *nodes, const double *props,
                                     const double *U, double *F, int stride) {
    double gp[4][3];
    gauss_points(gp);
    for (int e = 0; e < M; ++e) {
        double X[10][3], u[30], f[30] = {0}, C[6][6];
        for (int i = 0; i < 10; ++i) {
            const int n = conn[e * 10 + i];
            for (int c = 0; c < 3; ++c) { X[i][c] = nodes[3 * n + c]; u[3 * i + c] = U[(size_t)n * stride + c]; }
        }
        elasticity(props[2 * e], props[2 * e + 1], C);
        for (int g = 0; g < 4; ++g) {
            double B[6][30], detJ, eps[6] = {0}, sig[6] = {0};
            bmatrix(X, gp[g], B, &detJ);
            if (fabs(detJ) < 1e-15) continue;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 30; ++c) eps[r] += B[r][c] * u[c];
            for (int r = 0; r < 6; ++r)
                for (int k = 0; k < 6; ++k) sig[r] += C[r][k] * eps[k];
            for (int c = 0; c < 30; ++c) {
                double s = 0.0;
                for (int r = 0; r < 6; ++r) s += B[r][c] * sig[r];
                f[c] += OR_W * s * detJ;
            }
        }
        for (int i = 0; i < 10; ++i) {
            const int n = conn[e * 10 + i];
            for (int c = 0; c < 3; ++c) F[(size_t)n * stride + c] += f[3 * i + c];
        }
    }
}

static int find_col(const int *idx, int lo, int hi, int c) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (idx[mid] < c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* Accumulate element matrices into a CSR whose (sorted) pattern is given, in element
 * order then row-major 30x30 order: the duplicate-summation order of the reference's
 * COO (solid_solver.py:62-84). Serial => deterministic. */
OR_EXPORT int or_csr_accumulate(int M, const int *conn, const double *Ke, const int *indptr, const int *indices,
                                double *data) {
    for (int e = 0; e < M; ++e) {
        int dof[30];
        for (int i = 0; i < 10; ++i)
            for (int c = 0; c < 3; ++c) dof[3 * i + c] = 3 * conn[e * 10 + i] + c;
        const double *k = Ke + (size_t)e * 900;
        for (int a = 0; a < 30; ++a) {
            const int r = dof[a];
            for (int b = 0; b < 30; ++b) {
                const int p = find_col(indices, indptr[r], indptr[r + 1], dof[b]);
                if (p >= indptr[r + 1] || indices[p] != dof[b]) return -1;
                data[p] += k[a * 30 + b];
            }
        }
    }
    return 0;
}

OR_EXPORT void or_csr_spmv(int n, const int *indptr, const int *indices, const double *data, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int p = indptr[i]; p < indptr[i + 1]; ++p) s += data[p] * x[indices[p]];
        y[i] = s;
    }
}

static double dotd(int n, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Jacobi-preconditioned CG (FP64); diag_inv = 1/diag (or 1 where |diag|<=1e-12), x0 = 0,
 * stop when ||r|| <= rtol*||b||. Returns iterations; *rel = ||b-Kx||/||b|| recomputed.
 * solid_solver.py:228-236,302-308. */
OR_EXPORT int or_pcg(int n, const int *indptr, const int *indices, const double *data, const double *b, double *x,
                     double rtol, int max_iter, int use_jacobi, double *rel) {
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n), *dinv = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (indices[k] == i) d = data[k];
        dinv[i] = (use_jacobi && fabs(d) > 1e-12) ? 1.0 / d : 1.0;
        x[i] = 0.0; r[i] = b[i];
    }
    const double nb = sqrt(dotd(n, b, b));
    int it = 0;
    if (nb == 0.0) { *rel = 0.0; goto done; }
    for (int i = 0; i < n; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    double rho = dotd(n, r, z);
    for (it = 0; it < max_iter;) {
        or_csr_spmv(n, indptr, indices, data, p, q);
        const double alpha = rho / dotd(n, p, q);
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        ++it;
        if (sqrt(dotd(n, r, r)) <= rtol * nb) break;
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        const double rho1 = dotd(n, r, z);
        const double beta = rho1 / rho;
        rho = rho1;
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    or_csr_spmv(n, indptr, indices, data, x, q);
    for (int i = 0; i < n; ++i) q[i] = b[i] - q[i];
    *rel = sqrt(dotd(n, q, q)) / nb;
done:
    free(r); free(z); free(p); free(q); free(dinv);
    return it;
}

/* Mixed-precision iterative refinement (solid_solver.py:271-300): FP32 Jacobi-CG on a
 * float copy of K with tolerance max(1e-4, rtol) and max_iter/5 iterations, FP64 residual,
 * at most 10 outer steps. Returns total inner iterations. */
OR_EXPORT int or_ir_mixed(int n, const int *indptr, const int *indices, const double *data, const double *b, double *x,
                          double rtol, int max_iter, double *rel) {
    const int nnz = indptr[n];
    float *d32 = malloc(sizeof(float) * nnz), *dinv = malloc(sizeof(float) * n);
    float *r32 = malloc(sizeof(float) * n), *z = malloc(sizeof(float) * n), *p = malloc(sizeof(float) * n);
    float *q = malloc(sizeof(float) * n), *dx = malloc(sizeof(float) * n);
    double *r = malloc(sizeof(double) * n), *Kx = malloc(sizeof(double) * n);
    for (int k = 0; k < nnz; ++k) d32[k] = (float)data[k];
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (indices[k] == i) d = data[k];
        dinv[i] = (float)(fabs(d) > 1e-12 ? 1.0 / d : 1.0);
        x[i] = 0.0; r[i] = b[i];
    }
    const double nb = sqrt(dotd(n, b, b));
    int total = 0;
    const double itol = rtol > 1e-4 ? rtol : 1e-4;
    const int imax = max_iter / 5;
    if (nb == 0.0) { *rel = 0.0; goto out; }
    for (int outer = 0; outer < 10; ++outer) {
        if (sqrt(dotd(n, r, r)) / nb < rtol) break;
        double nr = 0.0;
        for (int i = 0; i < n; ++i) { r32[i] = (float)r[i]; dx[i] = 0.0f; nr += (double)r32[i] * r32[i]; }
        nr = sqrt(nr);
        for (int i = 0; i < n; ++i) { z[i] = dinv[i] * r32[i]; p[i] = z[i]; }
        float rho = 0.0f;
        for (int i = 0; i < n; ++i) rho += r32[i] * z[i];
        for (int it = 0; it < imax; ++it) {
            float pq = 0.0f;
            for (int i = 0; i < n; ++i) {
                float s = 0.0f;
                for (int k = indptr[i]; k < indptr[i + 1]; ++k) s += d32[k] * p[indices[k]];
                q[i] = s;
                pq += p[i] * s;
            }
            const float alpha = rho / pq;
            float rr = 0.0f;
            for (int i = 0; i < n; ++i) { dx[i] += alpha * p[i]; r32[i] -= alpha * q[i]; rr += r32[i] * r32[i]; }
            ++total;
            if (sqrtf(rr) <= itol * nr) break;
            float rho1 = 0.0f;
            for (int i = 0; i < n; ++i) { z[i] = dinv[i] * r32[i]; rho1 += r32[i] * z[i]; }
            const float beta = rho1 / rho;
            rho = rho1;
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        for (int i = 0; i < n; ++i) x[i] += (double)dx[i];
        or_csr_spmv(n, indptr, indices, data, x, Kx);
        for (int i = 0; i < n; ++i) r[i] = b[i] - Kx[i];
    }
    *rel = sqrt(dotd(n, r, r)) / nb;
out:
    free(d32); free(dinv); free(r32); free(z); free(p); free(q); free(dx); free(r); free(Kx);
    return total;
}

/* Reverse Cuthill-McKee of scipy.sparse.csgraph (symmetric_mode=True) on a CSR graph with
 * sorted column indices.  `seed_order` is numpy's argsort of the node degrees, where the
 * degree of row i is its entry count plus one if the diagonal is present.  Writes the
 * final (already reversed) order: perm[new] = old. */
OR_EXPORT void or_rcm(int n, const int *ptr, const int *ind, const int *seed_order, int *perm) {
    int *deg = malloc(sizeof(int) * n), *inds = malloc(sizeof(int) * n), *rev = malloc(sizeof(int) * n);
    int *order = malloc(sizeof(int) * (n > 0 ? n : 1));
    int maxdeg = 1;
    for (int i = 0; i < n; ++i) {
        deg[i] = ptr[i + 1] - ptr[i];
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (ind[k] == i) { deg[i] += 1; break; }
        if (deg[i] > maxdeg) maxdeg = deg[i];
    }
    int *tmp = malloc(sizeof(int) * (maxdeg + n + 1));
    for (int i = 0; i < n; ++i) { inds[i] = seed_order[i]; }
    for (int i = 0; i < n; ++i) rev[inds[i]] = i;
    int N = 0;
    for (int zz = 0; zz < n; ++zz) {
        if (inds[zz] != -1) {
            const int seed = inds[zz];
            order[N++] = seed;
            inds[rev[seed]] = -1;
            int level_start = N - 1, level_end = N;
            while (level_start < level_end) {
                for (int ii = level_start; ii < level_end; ++ii) {
                    const int i = order[ii];
                    const int N_old = N;
                    for (int jj = ptr[i]; jj < ptr[i + 1]; ++jj) {
                        const int j = ind[jj];
                        if (inds[rev[j]] != -1) {
                            inds[rev[j]] = -1;
                            order[N++] = j;
                        }
                    }
                    int level_len = 0;
                    for (int kk = N_old; kk < N; ++kk) tmp[level_len++] = deg[order[kk]];
                    for (int kk = 1; kk < level_len; ++kk) {
                        const int t = tmp[kk], t2 = order[N_old + kk];
                        int ll = kk;
                        while (ll > 0 && t < tmp[ll - 1]) {
                            tmp[ll] = tmp[ll - 1];
                            order[N_old + ll] = order[N_old + ll - 1];
                            --ll;
                        }
                        tmp[ll] = t;
                        order[N_old + ll] = t2;
                    }
                }
                level_start = level_end;
                level_end = N;
            }
        }
        if (N == n) break;
    }
    for (int i = 0; i < n; ++i) perm[i] = order[n - 1 - i];
    free(deg); free(inds); free(rev); free(order); free(tmp);
}

/* Cell-average stresses, cell-average von Mises, nodal average of cell VM
 * (solid_solver.py:311-365). u is 3 DOF/node. Returns first bad element + 1 or 0. */
OR_EXPORT int or_stress_field(int n_nodes, int M, const int *conn, const double *nodes, const double *E,
                              const double *nu, const double *u, double *cell_sig, double *cell_vm, double *nodal_vm) {
    int bad = 0;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < M; ++e) {
        double X[30], ue[30], sig[24], vm[4];
        gather_elem(conn, nodes, e, X);
        for (int i = 0; i < 10; ++i)
            for (int c = 0; c < 3; ++c) ue[3 * i + c] = u[3 * conn[e * 10 + i] + c];
        if (or_c3d10_stress(X, ue, E[e], nu[e], sig, vm)) {
#pragma omp critical
            if (!bad) bad = e + 1;
            continue;
        }
        for (int r = 0; r < 6; ++r) cell_sig[e * 6 + r] = (sig[r] + sig[6 + r] + sig[12 + r] + sig[18 + r]) / 4.0;
        cell_vm[e] = (vm[0] + vm[1] + vm[2] + vm[3]) / 4.0;
    }
    if (bad) return bad;
    double *acc = calloc(n_nodes, sizeof(double));
    int *cnt = calloc(n_nodes, sizeof(int));
    for (int e = 0; e < M; ++e)
        for (int i = 0; i < 10; ++i) { acc[conn[e * 10 + i]] += cell_vm[e]; cnt[conn[e * 10 + i]] += 1; }
    for (int i = 0; i < n_nodes; ++i) nodal_vm[i] = cnt[i] ? acc[i] / cnt[i] : 0.0;
    free(acc); free(cnt);
    return 0;
}
