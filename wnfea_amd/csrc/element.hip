// C3D10 element kernels for gfx950 (FP64).
//
// Element formulation follows wnfea/elements/c3d10.py: Abaqus node order
// (corners 1-4, midsides 1-2, 2-3, 3-1, 1-4, 2-4, 3-4; :35-106), 4-point
// Hammer rule a = (5-sqrt5)/20, b = (5+3sqrt5)/20, w = 1/24 (:22-32), isotropic
// D (:192-216), Ke = sum_g w B^T D B det J with DOF order 3*node+comp
// (:219-249), f_int = sum_g w B^T sigma det J (:252-283), Gauss-point stress
// and von Mises (:286-318).
//
// Mapping: one 64-lane wavefront per element.  Lanes 0..39 each own one
// (Gauss point, node) pair and compute that node's Cartesian gradient g_a at
// that point (the 3x3 Jacobian is recomputed per lane, 30 FMAs, instead of a
// cross-lane broadcast).  The isotropic block formula
//   K_ab = sum_g w detJ (lam g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I)
// is algebraically B_a^T D B_b, so each of the 55 lanes owning an upper
// block pair (a <= b) produces a full 3x3 block and also writes its transpose:
// Ke comes out exactly symmetric, as the reference's 0.5 (Ke + Ke^T).
#include <algorithm>
#include <type_traits>
#include <utility>
#include "common.hpp"
#include "wn.hpp"

namespace wn {

__constant__ double c_gauss[4][3];
__constant__ double c_dndxi[4][3][10];  // natural derivatives per Gauss point

struct ElemConsts {
  double gauss[4][3];
  double dndxi[4][3][10];
};

static ElemConsts make_consts() {
  ElemConsts c{};
  const double a = (5.0 - sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
  const double pts[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
  for (int g = 0; g < 4; ++g) {
    for (int k = 0; k < 3; ++k) c.gauss[g][k] = pts[g][k];
    const double L1 = 1.0 - pts[g][0] - pts[g][1] - pts[g][2], L2 = pts[g][0], L3 = pts[g][1], L4 = pts[g][2];
    // dN_i/dL_k for the 10 shape functions, then d/dxi_m = d/dL_{m+1} - d/dL_1
    double dl[4][10] = {};
    dl[0][0] = 4 * L1 - 1; dl[1][1] = 4 * L2 - 1; dl[2][2] = 4 * L3 - 1; dl[3][3] = 4 * L4 - 1;
    dl[0][4] = 4 * L2; dl[1][4] = 4 * L1;
    dl[1][5] = 4 * L3; dl[2][5] = 4 * L2;
    dl[0][6] = 4 * L3; dl[2][6] = 4 * L1;
    dl[0][7] = 4 * L4; dl[3][7] = 4 * L1;
    dl[1][8] = 4 * L4; dl[3][8] = 4 * L2;
    dl[2][9] = 4 * L4; dl[3][9] = 4 * L3;
    for (int m = 0; m < 3; ++m)
      for (int i = 0; i < 10; ++i) c.dndxi[g][m][i] = dl[m + 1][i] - dl[0][i];
  }
  return c;
}

// Uploads the quadrature tables to the current device once per process
// (not thread-safe; call from the thread that owns the device).
int upload_element_consts(hipStream_t s) {
  static unsigned done_mask = 0;
  int dev = 0;
  WN_HIP(hipGetDevice(&dev));
  if (dev < 32 && ((done_mask >> dev) & 1u)) return kOk;
  const ElemConsts c = make_consts();
  WN_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_gauss), c.gauss, sizeof(c.gauss), 0, hipMemcpyHostToDevice, s));
  WN_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_dndxi), c.dndxi, sizeof(c.dndxi), 0, hipMemcpyHostToDevice, s));
  WN_HIP(hipStreamSynchronize(s));
  if (dev < 32) done_mask |= 1u << dev;
  return kOk;
}

// Cartesian gradient of node `a` at Gauss point `g`; returns det J.
__device__ __forceinline__ double node_gradient(const double (*X)[3], int g, int a, double grad[3]) {
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 10; ++i) s = fma(c_dndxi[g][r][i], X[i][c], s);
      J[r][c] = s;
    }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double id = 1.0 / det;
  const double inv[3][3] = {
      {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
      {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
      {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
  const double d0 = c_dndxi[g][0][a], d1 = c_dndxi[g][1][a], d2 = c_dndxi[g][2][a];
#pragma unroll
  for (int r = 0; r < 3; ++r) grad[r] = inv[r][0] * d0 + inv[r][1] * d1 + inv[r][2] * d2;
  return det;
}

constexpr int kElemPerBlock = 4;  // 4 wavefronts, one element each

// Ke[e][30][30] (row-major, DOF 3*node+comp).  bad_flag gets (e+1) of an element
// with det J <= 0 (the smallest such index is not guaranteed; any one is reported).
__global__ __launch_bounds__(256) void c3d10_ke_kernel(int M, const int* __restrict__ conn,
                                                       const double* __restrict__ nodes,
                                                       const double* __restrict__ E,
                                                       const double* __restrict__ nu, double* __restrict__ Ke,
                                                       int* __restrict__ bad_flag) {
  __shared__ double sX[kElemPerBlock][10][3];
  __shared__ double sG[kElemPerBlock][4][10][3];  // w detJ folded in separately
  __shared__ double sW[kElemPerBlock][4];
  __shared__ __align__(16) double sK[kElemPerBlock][900];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e = blockIdx.x * kElemPerBlock + w;
  const bool live = e < M;
  if (live && lane < 30) {
    const int node = conn[(size_t)e * 10 + lane / 3];
    sX[w][lane / 3][lane % 3] = nodes[(size_t)node * 3 + lane % 3];
  }
  __syncthreads();
  if (live && lane < 40) {
    const int g = lane / 10, a = lane % 10;
    double grad[3];
    const double det = node_gradient(sX[w], g, a, grad);
    sG[w][g][a][0] = grad[0];
    sG[w][g][a][1] = grad[1];
    sG[w][g][a][2] = grad[2];
    if (a == 0) {
      sW[w][g] = det / 24.0;
      if (!(det > 0.0)) atomicMax(bad_flag, e + 1);
    }
  }
  __syncthreads();
  // lane -> upper-triangle pair (a, b), a <= b; the 55 pair blocks (and their
  // mirror images) are assembled in LDS, then the wavefront stores its
  // element's 900 values as contiguous 16-byte chunks (a lane's own 3x3 blocks
  // would be 24-byte pieces scattered over 30 rows)
  if (live && lane < 55) {
    int a = 0, rem = lane;
    while (rem >= 10 - a) { rem -= 10 - a; ++a; }
    const int b = a + rem;
    const double Ev = E[e], nv = nu[e];
    const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
    const double lam = f * nv, mu = Ev / (2.0 * (1.0 + nv));
    double K[3][3] = {};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const double wd = sW[w][g];
      const double ga[3] = {sG[w][g][a][0], sG[w][g][a][1], sG[w][g][a][2]};
      const double gb[3] = {sG[w][g][b][0], sG[w][g][b][1], sG[w][g][b][2]};
      const double dot = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double v = lam * ga[r] * gb[c] + mu * ga[c] * gb[r];
          if (r == c) v += mu * dot;
          K[r][c] = fma(wd, v, K[r][c]);
        }
    }
    double* out = sK[w];
    if (a == b) {  // diagonal block: mirror so that rounding cannot break symmetry
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) out[(3 * a + r) * 30 + 3 * a + c] = 0.5 * (K[r][c] + K[c][r]);
    } else {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          out[(3 * a + r) * 30 + 3 * b + c] = K[r][c];
          out[(3 * b + c) * 30 + 3 * a + r] = K[r][c];
        }
    }
  }
  __syncthreads();
  if (!live) return;
  // 900 doubles = 450 16-byte chunks; Ke + 900 e is 16-byte aligned
  using D2 = double __attribute__((ext_vector_type(2)));
  const D2* src = reinterpret_cast<const D2*>(sK[w]);
  D2* dst = reinterpret_cast<D2*>(Ke + (size_t)e * 900);
  for (int k = lane; k < 450; k += 64) __builtin_nontemporal_store(src[k], dst + k);
}

// The same Ke with the contraction on the matrix cores:
//   Ke = sum_g w_g B_g^T D B_g = A Bd,  A[i][k] = B_g[s][i],  Bd[k][j] = w_g (D B_g)[s][j],
// k = 6 g + s over the 4 Gauss points and the 6 strain components (xx, yy, zz,
// xy, yz, xz; engineering shears), i, j the 30 DOFs padded to 32.  One
// wavefront per element: the operands A (32 x 24) and Bd (24 x 32) are built
// in LDS by the whole wavefront, then v_mfma_f64_16x16x4f64 runs on the three
// 16x16 tiles (0,0), (0,1), (1,1) of the upper triangle, six K-steps of 4
// each.  Operand and result lanes (checked on the MI355X): lane l supplies
// A[16 I + l % 16][4 t + l / 16] and Bd[4 t + l / 16][16 J + l % 16], and
// result r of lane l is D[4 r + l / 16][l % 16].  The results overwrite the
// operands' LDS; the tile (0,1) is mirrored into (1,0), and the diagonal
// tiles are symmetrised (0.5 (K + K^T), as the reference does) while the
// element is stored.
typedef double mf_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double strain_b(const double* gr, int s, int c) {
  // B_g[s][3 a + c] for the node gradient gr = (gx, gy, gz)
  switch (s) {
    case 0: return c == 0 ? gr[0] : 0.0;
    case 1: return c == 1 ? gr[1] : 0.0;
    case 2: return c == 2 ? gr[2] : 0.0;
    case 3: return c == 0 ? gr[1] : (c == 1 ? gr[0] : 0.0);
    case 4: return c == 1 ? gr[2] : (c == 2 ? gr[1] : 0.0);
    default: return c == 0 ? gr[2] : (c == 2 ? gr[0] : 0.0);
  }
}

// A (32 rows x 24, row stride 25) then Bd (24 rows x 32, row stride 33): the
// odd strides spread a wavefront's operand reads and transposed writes over
// the LDS banks.  Two elements (wavefronts) per block keep the LDS per block
// at 27 KB.
constexpr int kKeAS = 25, kKeBS = 33;
constexpr int kKeOps = 32 * kKeAS + 24 * kKeBS;
constexpr int kKeMfElem = 2;

__global__ __launch_bounds__(64 * kKeMfElem) void c3d10_ke_mfma_kernel(int M, const int* __restrict__ conn,
                                                            const double* __restrict__ nodes,
                                                            const double* __restrict__ E,
                                                            const double* __restrict__ nu, double* __restrict__ Ke,
                                                            int* __restrict__ bad_flag) {
  __shared__ double sX[kKeMfElem][10][3];
  __shared__ double sG[kKeMfElem][4][10][3];
  __shared__ double sW[kKeMfElem][4];
  __shared__ __align__(16) double sOp[kKeMfElem][kKeOps];  // operands, then the 32 x 32 result
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e = blockIdx.x * kKeMfElem + w;
  const bool live = e < M;
  if (live && lane < 30) {
    const int node = conn[(size_t)e * 10 + lane / 3];
    sX[w][lane / 3][lane % 3] = nodes[(size_t)node * 3 + lane % 3];
  }
  __syncthreads();
  if (live && lane < 40) {
    const int g = lane / 10, a = lane % 10;
    double grad[3];
    const double det = node_gradient(sX[w], g, a, grad);
    sG[w][g][a][0] = grad[0];
    sG[w][g][a][1] = grad[1];
    sG[w][g][a][2] = grad[2];
    if (a == 0) {
      sW[w][g] = det / 24.0;
      if (!(det > 0.0)) atomicMax(bad_flag, e + 1);
    }
  }
  __syncthreads();
  double* A = sOp[w];
  double* Bd = sOp[w] + 32 * kKeAS;
  if (live) {
    const double Ev = E[e], nv = nu[e];
    const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
    const double lam = f * nv, mu = Ev / (2.0 * (1.0 + nv)), l2m = lam + 2.0 * mu;
    // A[i][k] and Bd[k][i] for i = DOF, k = 6 g + s: 768 entries each, 12 per lane
    for (int q = lane; q < 32 * 24; q += 64) {
      const int i = q / 24, k = q % 24, g = k / 6, sc = k % 6;
      double av = 0.0, bv = 0.0;
      if (i < 30) {
        const double* gi = sG[w][g][i / 3];
        const int ci = i % 3;
        av = strain_b(gi, sc, ci);
        double db;
        if (sc < 3) {
          const double b0 = strain_b(gi, 0, ci), b1 = strain_b(gi, 1, ci), b2 = strain_b(gi, 2, ci);
          db = l2m * av + lam * (b0 + b1 + b2 - av);
        } else {
          db = mu * av;
        }
        bv = sW[w][g] * db;
      }
      A[i * kKeAS + k] = av;
      Bd[k * kKeBS + i] = bv;
    }
  }
  __syncthreads();
  mf_d4 acc[3];
  if (live) {
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int tile = 0; tile < 3; ++tile) {
      const int I = tile == 2 ? 1 : 0, J = tile == 0 ? 0 : 1;
      acc[tile] = mf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        const int k = 4 * t + lk;
        acc[tile] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(16 * I + lr) * kKeAS + k], Bd[k * kKeBS + 16 * J + lr],
                                                         acc[tile], 0, 0, 0);
      }
    }
  }
  __syncthreads();  // every lane is done with the operands: reuse their LDS for the result
  if (live) {
    const int lr = lane & 15, lk = lane >> 4;
    double* K2 = sOp[w];
#pragma unroll
    for (int tile = 0; tile < 3; ++tile) {
      const int I = tile == 2 ? 1 : 0, J = tile == 0 ? 0 : 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * I + 4 * r + lk, col = 16 * J + lr;
        K2[row * 32 + col] = acc[tile][r];
        if (I != J) K2[col * 32 + row] = acc[tile][r];
      }
    }
  }
  __syncthreads();
  if (!live) return;
  // symmetrise and store the 30 x 30 element as 16-byte chunks
  using D2 = double __attribute__((ext_vector_type(2)));
  D2* dst = reinterpret_cast<D2*>(Ke + (size_t)e * 900);
  const double* k2 = sOp[w];
  for (int q = lane; q < 450; q += 64) {
    const int i0 = 2 * q, r0 = i0 / 30, c0 = i0 % 30;  // c0 even, so c0 + 1 < 30 is in the same row
    D2 v;
    v.x = 0.5 * (k2[r0 * 32 + c0] + k2[c0 * 32 + r0]);
    v.y = 0.5 * (k2[r0 * 32 + c0 + 1] + k2[(c0 + 1) * 32 + r0]);
    __builtin_nontemporal_store(v, dst + q);
  }
}

// Ke variant: 0 vector ALU (default), 1 matrix cores
static int g_ke_variant = 0;
int set_ke_variant(int v) {
  if (v != 0 && v != 1) return kBadArg;
  g_ke_variant = v;
  return kOk;
}

// Per-element internal forces fe[e][30] = Ke(e) u_e, matrix-free (no Ke), plus
// optional Gauss-point stresses.  One thread per element, FP64.  U is read with
// DOF stride `stride` (3 for solid vectors, 6 for the reference's mixed layout).
// A Gauss point with |det J| < 1e-15 is skipped (kernels.cpp:312 semantics).
__global__ __launch_bounds__(128) void c3d10_fint_elem_kernel(int M, const int* __restrict__ conn,
                                                              const double* __restrict__ nodes,
                                                              const double* __restrict__ E,
                                                              const double* __restrict__ nu,
                                                              const double* __restrict__ U, int stride,
                                                              double* __restrict__ fe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  double X[10][3], u[10][3], fo[10][3] = {};
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const int n = conn[(size_t)e * 10 + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      X[i][c] = nodes[(size_t)n * 3 + c];
      u[i][c] = U[(size_t)n * stride + c];
    }
  }
  const double Ev = E[e], nv = nu[e];
  const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
  const double c11 = f * (1.0 - nv), c12 = f * nv, c44 = Ev / (2.0 * (1.0 + nv));
  for (int g = 0; g < 4; ++g) {
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 10; ++i) s = fma(c_dndxi[g][r][i], X[i][c], s);
        J[r][c] = s;
      }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (fabs(det) < 1e-15) continue;
    const double id = 1.0 / det;
    const double inv[3][3] = {
        {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id, (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id},
        {c01 * id, (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id},
        {c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id, (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id}};
    double gr[10][3];
    double eps[6] = {};
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const double d0 = c_dndxi[g][0][i], d1 = c_dndxi[g][1][i], d2 = c_dndxi[g][2][i];
#pragma unroll
      for (int r = 0; r < 3; ++r) gr[i][r] = inv[r][0] * d0 + inv[r][1] * d1 + inv[r][2] * d2;
      eps[0] += gr[i][0] * u[i][0];
      eps[1] += gr[i][1] * u[i][1];
      eps[2] += gr[i][2] * u[i][2];
      eps[3] += gr[i][1] * u[i][0] + gr[i][0] * u[i][1];
      eps[4] += gr[i][2] * u[i][1] + gr[i][1] * u[i][2];
      eps[5] += gr[i][2] * u[i][0] + gr[i][0] * u[i][2];
    }
    const double sig[6] = {c11 * eps[0] + c12 * (eps[1] + eps[2]), c11 * eps[1] + c12 * (eps[0] + eps[2]),
                           c11 * eps[2] + c12 * (eps[0] + eps[1]), c44 * eps[3], c44 * eps[4], c44 * eps[5]};
    const double wd = det / 24.0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      fo[i][0] += wd * (gr[i][0] * sig[0] + gr[i][1] * sig[3] + gr[i][2] * sig[5]);
      fo[i][1] += wd * (gr[i][1] * sig[1] + gr[i][0] * sig[3] + gr[i][2] * sig[4]);
      fo[i][2] += wd * (gr[i][2] * sig[2] + gr[i][1] * sig[4] + gr[i][0] * sig[5]);
    }
  }
  double* o = fe + (size_t)e * 30;
#pragma unroll
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) o[3 * i + c] = fo[i][c];
}

// Deterministic node-wise gather of element vectors: out[n*stride + c] =
// (accumulate ? out : 0) + sum over the node's (element, local) entries in
// element order of fe[e][3*local + c].
__global__ void gather_node_kernel(int n_nodes, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                                   const double* __restrict__ fe, double* __restrict__ out, int stride,
                                   int accumulate) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; ++k) {
    const int code = n2e[k];  // e * 10 + local
    const double* p = fe + (size_t)(code / 10) * 30 + 3 * (code % 10);
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
  }
  double* o = out + (size_t)n * stride;
  if (accumulate) {
    s0 += o[0];
    s1 += o[1];
    s2 += o[2];
  }
  o[0] = s0;
  o[1] = s1;
  o[2] = s2;
}

// Gauss-point stresses -> cell averages (sigma, von Mises) per element; u has
// 3 DOF per node.  Mirrors solid_solver.py:339-355 (mean of the 4 Gauss-point
// stresses and of the 4 Gauss-point von Mises values).
__global__ __launch_bounds__(128) void c3d10_stress_kernel(int M, const int* __restrict__ conn,
                                                           const double* __restrict__ nodes,
                                                           const double* __restrict__ E,
                                                           const double* __restrict__ nu,
                                                           const double* __restrict__ U,
                                                           double* __restrict__ gp_sig,  // [M][4][6] or null
                                                           double* __restrict__ gp_vm,   // [M][4] or null
                                                           double* __restrict__ cell_sig,
                                                           double* __restrict__ cell_vm, int* __restrict__ bad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  double X[10][3], u[10][3];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const int n = conn[(size_t)e * 10 + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      X[i][c] = nodes[(size_t)n * 3 + c];
      u[i][c] = U[(size_t)n * 3 + c];
    }
  }
  const double Ev = E[e], nv = nu[e];
  const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
  const double c11 = f * (1.0 - nv), c12 = f * nv, c44 = Ev / (2.0 * (1.0 + nv));
  double sacc[6] = {}, vacc = 0.0;
  for (int g = 0; g < 4; ++g) {
    double eps[6] = {};
    double det = 0.0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      double gr[3];
      det = node_gradient(X, g, i, gr);
      eps[0] += gr[0] * u[i][0];
      eps[1] += gr[1] * u[i][1];
      eps[2] += gr[2] * u[i][2];
      eps[3] += gr[1] * u[i][0] + gr[0] * u[i][1];
      eps[4] += gr[2] * u[i][1] + gr[1] * u[i][2];
      eps[5] += gr[2] * u[i][0] + gr[0] * u[i][2];
    }
    if (!(det > 0.0)) atomicMax(bad, e + 1);
    const double s[6] = {c11 * eps[0] + c12 * (eps[1] + eps[2]), c11 * eps[1] + c12 * (eps[0] + eps[2]),
                         c11 * eps[2] + c12 * (eps[0] + eps[1]), c44 * eps[3], c44 * eps[4], c44 * eps[5]};
    const double vm = sqrt(0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) +
                                  (s[2] - s[0]) * (s[2] - s[0]) + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])));
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      sacc[r] += s[r];
      if (gp_sig) gp_sig[((size_t)e * 4 + g) * 6 + r] = s[r];
    }
    if (gp_vm) gp_vm[(size_t)e * 4 + g] = vm;
    vacc += vm;
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) cell_sig[(size_t)e * 6 + r] = sacc[r] / 4.0;
  cell_vm[e] = vacc / 4.0;
}

// Nodal average of cell von Mises over incident elements (deterministic gather).
__global__ void nodal_average_kernel(int n_nodes, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                                     const double* __restrict__ cell_vm, double* __restrict__ nodal_vm) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  const int b = n2e_ptr[n], t = n2e_ptr[n + 1];
  double s = 0.0;
  for (int k = b; k < t; ++k) s += cell_vm[n2e[k] / 10];
  nodal_vm[n] = (t > b) ? s / (t - b) : 0.0;
}

// Adjoint sensitivity of J(u) with dK/dE_e = K_e/E_e (K linear in E_e):
//   dJ/dE_e = -(1/E_e) lambda_e^T K_e u_e, with fe = K_e u_e from the f_int kernel.
__global__ void adjoint_dot_kernel(int M, const int* __restrict__ conn, const double* __restrict__ E,
                                   const double* __restrict__ fe, const double* __restrict__ lam,
                                   double* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const int n = conn[(size_t)e * 10 + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) s += lam[(size_t)n * 3 + c] * fe[(size_t)e * 30 + 3 * i + c];
  }
  grad[e] = -s / E[e];
}

// ---------------------------------------------------------------- matrix-free K u
// Quadrature data of the matrix-free operator ("partial assembly"): for each
// element and Gauss point g the inverse Jacobian J^-1 (row-major) and
// det J / 24, stored qd[(k*M + e)*4 + g] so that the four lanes of an
// element and the elements of a wavefront read one contiguous run.  A
// non-positive det J flags the element (bad = e + 1).
__global__ __launch_bounds__(128) void c3d10_qdata_kernel(int M, const int* __restrict__ conn,
                                                          const double* __restrict__ nodes, double* __restrict__ qd,
                                                          int* bad, int* curved) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  double X[10][3];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const int n = conn[(size_t)e * 10 + i];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[i][c] = nodes[(size_t)n * 3 + c];
  }
  double J0[9] = {};  // J at the first Gauss point
  bool differs = false;
  for (int g = 0; g < 4; ++g) {
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 10; ++i) s = fma(c_dndxi[g][r][i], X[i][c], s);
        J[r][c] = s;
      }
    // affine element: J equal at all points up to rounding
    double jmax = 0.0, dmax = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if (g == 0) J0[k] = J[k / 3][k % 3];
      jmax = fmax(jmax, fabs(J0[k]));
      dmax = fmax(dmax, fabs(J[k / 3][k % 3] - J0[k]));
    }
    differs = differs || dmax > 1e-12 * jmax;
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (!(det > 0.0) && bad) atomicCAS(bad, 0, e + 1);
    const double id = det != 0.0 ? 1.0 / det : 0.0;
    const double inv[9] = {c00 * id, (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id,
                           (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id, c01 * id,
                           (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id, (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id,
                           c02 * id, (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id,
                           (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id};
#pragma unroll
    for (int k = 0; k < 9; ++k) qd[((size_t)k * M + e) * 4 + g] = inv[k];
    qd[((size_t)9 * M + e) * 4 + g] = det / 24.0;
  }
  if (differs && curved) atomicOr(curved, 1);
}

// Once-per-element quadrature data of an affine mesh: the first Gauss
// point's entries, qa[k*M + e] = qd[(k*M + e)*4].
__global__ void qdata_affine_kernel(long long n, const double* __restrict__ qd, double* __restrict__ qa) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    qa[i] = qd[i * 4];
}

// v from the lane given by a quad permutation (DPP, within groups of 4
// lanes; every lane reads a lane of its own quad, so all are written).  The
// FP32 form has no old value (bound_ctrl set), which lets the compiler fold
// the permutation into the consuming add or multiply (v_add_f32_dpp) without
// first clearing the destination.
template <int Ctrl>
__device__ __forceinline__ float quad_perm(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), Ctrl, 0xF, 0xF, true));
}
template <int Ctrl>
__device__ __forceinline__ double quad_perm(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(u & 0xffffffffu), Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), Ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// v of lane o (a compile-time constant after unrolling) of the quad
template <int O>
__device__ __forceinline__ float quad_bcast_c(float v) {
  return quad_perm<O | (O << 2) | (O << 4) | (O << 6)>(v);
}
template <int O>
__device__ __forceinline__ double quad_bcast_c(double v) {
  return quad_perm<O | (O << 2) | (O << 4) | (O << 6)>(v);
}
template <typename T>
__device__ __forceinline__ T quad_bcast(T v, int o) {
  switch (o) {
    case 0: return quad_bcast_c<0>(v);
    case 1: return quad_bcast_c<1>(v);
    case 2: return quad_bcast_c<2>(v);
    default: return quad_bcast_c<3>(v);
  }
}

// Sum over the 4 lanes of a quad; every lane gets the same bits,
// (v0 + v1) + (v2 + v3) up to commutation.  Plain adds (no contraction with
// the product that made v), so each step is one v_add_f32_dpp in FP32.
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
#pragma clang fp contract(off)
  v = quad_perm<0xB1>(v) + v;  // lanes (1, 0, 3, 2): xor 1
  v = quad_perm<0x4E>(v) + v;  // lanes (2, 3, 0, 1): xor 2
  return v;
}

// The 10-node tetrahedron in barycentric derivatives: node i's shape function
// depends on L_k only through corner k (N = L_k (2 L_k - 1), dN/dL_k =
// 4 L_k - 1) or through the two ends (p, q) of its edge (N = 4 L_p L_q,
// dN/dL_p = 4 L_q), and d/dxi_a = d/dL_{a+1} - d/dL_0.  Midside nodes 4..9
// sit on the edges 1-2, 2-3, 3-1, 1-4, 2-4, 3-4 (Abaqus order).
__device__ constexpr int kEdgeP[6] = {0, 1, 2, 0, 1, 2};
__device__ constexpr int kEdgeQ[6] = {1, 2, 0, 3, 3, 3};

// The element forces of one quad of lanes (lane g: Gauss point g; see
// c3d10_pa_elem_kernel): from node i's scaled values (get(i, u) fills u[3]),
// the point's inverse Jacobian Ji and the weighted moduli w11 = w c11,
// w12 = w c12, w44 = w c44 (w = det J / 24), the quad-summed forces of the
// lane's nodes g, g + 4, g + 8 (out[j] for node g + 4 j; out[2] stays 0
// for g >= 2).
template <typename T, typename Get>
__device__ __forceinline__ void pa_forces(int g, const T (&c4)[4], const T (&cc)[4], const T (&Ji)[9], T w11, T w12,
                                          T w44, Get get, T (&out)[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[j][c] = T(0);
  T Q[4][3];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    T u[3];
    get(i, u);
    if (i < 4) {
#pragma unroll
      for (int c = 0; c < 3; ++c) Q[i][c] = cc[i] * u[c];  // corners come first: initialise
    } else {
      const int p = kEdgeP[i - 4], q = kEdgeQ[i - 4];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Q[p][c] += c4[q] * u[c];
        Q[q][c] += c4[p] * u[c];
      }
    }
  }
  T H[3][3];  // H = J^-1 G,  G[a][c] = Q[a+1][c] - Q[0][c]
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T g0 = Q[1][c] - Q[0][c], g1 = Q[2][c] - Q[0][c], g2 = Q[3][c] - Q[0][c];
#pragma unroll
    for (int r = 0; r < 3; ++r) H[r][c] = Ji[3 * r] * g0 + Ji[3 * r + 1] * g1 + Ji[3 * r + 2] * g2;
  }
  const T e0 = H[0][0], e1 = H[1][1], e2 = H[2][2];
  const T s0 = w11 * e0 + w12 * (e1 + e2);
  const T s1 = w11 * e1 + w12 * (e0 + e2);
  const T s2 = w11 * e2 + w12 * (e0 + e1);
  const T s3 = w44 * (H[1][0] + H[0][1]), s4 = w44 * (H[2][1] + H[1][2]), s5 = w44 * (H[2][0] + H[0][2]);
  // S = [[s0 s3 s5] [s3 s1 s4] [s5 s4 s2]];  R_{k+1} = column k of S J^-1, R_0 = -(R_1 + R_2 + R_3)
  T R[4][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    R[k + 1][0] = s0 * Ji[k] + s3 * Ji[3 + k] + s5 * Ji[6 + k];
    R[k + 1][1] = s3 * Ji[k] + s1 * Ji[3 + k] + s4 * Ji[6 + k];
    R[k + 1][2] = s5 * Ji[k] + s4 * Ji[3 + k] + s2 * Ji[6 + k];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) R[0][c] = -(R[1][c] + R[2][c] + R[3][c]);
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    T fi[3];
    if (i < 4) {
#pragma unroll
      for (int c = 0; c < 3; ++c) fi[c] = R[i][c] * cc[i];
    } else {
      const int p = kEdgeP[i - 4], q = kEdgeQ[i - 4];
#pragma unroll
      for (int c = 0; c < 3; ++c) fi[c] = R[p][c] * c4[q] + R[q][c] * c4[p];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) fi[c] = quad_sum(fi[c]);
    if ((i & 3) == g) {
#pragma unroll
      for (int c = 0; c < 3; ++c) out[i >> 2][c] = fi[c];
    }
  }
}

// Weighted isotropic moduli (w c11, w c12, w c44) from E, nu (in FP64, rounded to T).
template <typename T>
__device__ __forceinline__ void weighted_moduli(T wd, double Ev, double nv, T& w11, T& w12, T& w44) {
  const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
  w11 = wd * T(f * (1.0 - nv));
  w12 = wd * T(f * nv);
  w44 = wd * T(Ev / (2.0 * (1.0 + nv)));
}

// 4 L_k and 4 L_k - 1 at Gauss point g of the 4-point rule: L_k = b for
// k = g, else a, with a = (5 - sqrt 5) / 20 and b = (5 + 3 sqrt 5) / 20
// (= 1 - 3 a); the products with 4 are exact, so every constant is one
// selection between two literals.
template <typename T>
__device__ __forceinline__ void gauss_point_consts(int g, T (&c4)[4], T (&cc)[4]) {
  constexpr double ga = 0.13819660112501051517954131656343619;  // (5 - sqrt 5) / 20
  constexpr double gb = 0.58541019662496845446137605030969143;  // (5 + 3 sqrt 5) / 20
  constexpr T a4 = T(4.0 * ga), b4 = T(4.0 * gb), a4m = T(4.0 * ga) - T(1), b4m = T(4.0 * gb) - T(1);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    c4[k] = k == g ? b4 : a4;
    cc[k] = k == g ? b4m : a4m;
  }
}

// fe = K_e (sc .* u)_e from the quadrature data, stored [i][e][3] for local
// node i, where sc is a per-DOF scaling that also carries the Dirichlet rule
// (0 on a fixed DOF: its column of the operator is zero).  Four lanes per
// element, lane g for Gauss point g.  Strain: Q_k = sum_i dN_i/dL_k u_i (only
// the nonzero derivatives are visited), the natural displacement gradient
// G_a = Q_{a+1} - Q_0 and H = J^-1 G, H[r][c] = du_c/dx_r.  Stress S = D eps,
// T = (det J / 24) S J^-1, and node i's force T dN_i/dxi = sum_k R_k dN_i/dL_k
// with R_k = T e_{k-1} (k >= 1), R_0 = -(R_1 + R_2 + R_3); the four points'
// forces are summed over the quad and lane g stores the nodes i = g (mod 4).
// The four-fold split keeps enough waves in flight to hide the latency of the
// connectivity and nodal-value loads.  T is the type of the quadrature data,
// the vectors and the arithmetic; Affine: one set of quadrature data per
// element instead of one per Gauss point.
template <typename T, bool Affine>
__global__ __launch_bounds__(256) void c3d10_pa_elem_kernel(int M, const int* __restrict__ conn,
                                                            const T* __restrict__ qd,
                                                            const double* __restrict__ E,
                                                            const double* __restrict__ nu,
                                                            const T* __restrict__ sc,
                                                            const T* __restrict__ U, T* __restrict__ fe,
                                                            const int* flags, int guard) {
  if (guard && flags[0]) return;
  // XCD-contiguous element ranges (the grid is padded to a multiple of 8)
  const int nblk = (int)((4LL * M + 255) / 256);
  const int lb = slice_of_block(nblk, 1);
  if (lb >= nblk) return;
  const long long t = (long long)lb * blockDim.x + threadIdx.x;
  const int e = (int)(t >> 2), g = (int)(t & 3);
  if (e >= M) return;  // whole quads leave together
  T c4[4], cc[4];  // 4 L_k and 4 L_k - 1
  gauss_point_consts<T>(g, c4, cc);
  // Loads in two rounds, each issued together: first everything that does
  // not depend on another load (the connectivity of this lane's nodes g,
  // g + 4, g + 8, the quadrature data, the moduli), then the nodal values
  // the connectivity addresses.  A lane without a third node (g >= 2) reads
  // node 9's entries and drops them, which keeps the code free of branches
  // (left to itself the compiler serialised the third node's chain and the
  // moduli behind the rest: four exposed memory latencies instead of two).
  int nd[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) nd[j] = conn[(size_t)e * 10 + (g + 4 * j < 10 ? g + 4 * j : 9)];
  // quadrature data of this point (affine: of the element, shared by its 4 lanes)
  T Ji[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Ji[k] = Affine ? qd[(size_t)k * M + e] : qd[((size_t)k * M + e) * 4 + g];
  const T wd = Affine ? qd[(size_t)9 * M + e] : qd[((size_t)9 * M + e) * 4 + g];
  const double Ev = E[e], nv = nu[e];
  __builtin_amdgcn_sched_group_barrier(0x020, 15, 0);
  // lane g loads the scaled values of nodes g, g + 4, g + 8 only; the quad
  // shares them by DPP broadcast (a third of the memory instructions)
  T own[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int n3 = 3 * nd[j];
    const T keep = g + 4 * j < 10 ? T(1) : T(0);
#pragma unroll
    for (int c = 0; c < 3; ++c) own[j][c] = keep * (U[n3 + c] * sc[n3 + c]);
  }
  __builtin_amdgcn_sched_group_barrier(0x020, 18, 0);
  T w11, w12, w44;
  weighted_moduli<T>(wd, Ev, nv, w11, w12, w44);
  T out[3][3];
  pa_forces<T>(g, c4, cc, Ji, w11, w12, w44,
               [&](int i, T (&u)[3]) {
#pragma unroll
                 for (int c = 0; c < 3; ++c) u[c] = quad_bcast(own[i >> 2][c], i & 3);
               },
               out);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int i = g + 4 * j;
    if (i < 10) {
      // layout [local node][element][3]
      T* o = fe + ((size_t)i * M + e) * 3;
      o[0] = out[j][0];
      o[1] = out[j][1];
      o[2] = out[j][2];
    }
  }
}

// Node-wise deterministic gather with the assembled operator's Dirichlet rule
// (a fixed DOF has a unit diagonal and no coupling): for every DOF d of node n
//   mode 0: y_d = (K u)_d;   mode 1: y_d = (K u)_d and partial u.y;
//   mode 2: y_d = f_d - (K u)_d and partial y.y,
// where (K u)_d = u_d for a fixed DOF, else the sum of the node's element
// contributions in element order.  Each wavefront takes one chunk of whole
// nodes, at most 64 of them with at most kChunkInc incidences in all (chunk
// boundaries fixed at set-up): lane l loads incidences l, l + 64, ... of the
// chunk (one round of independent loads, issued together with the node's own
// operands), the values meet in LDS, and one lane per node adds its
// incidences in order.  A node with more than kChunkInc incidences is a chunk
// of its own, summed by its first lane from memory.
constexpr int kGatherWaves = 4;
constexpr int kIncPerLane = 4;
constexpr int kChunkInc = 64 * kIncPerLane;

__global__ __launch_bounds__(64 * kGatherWaves) void mf_gather_kernel(
    int n_chunks, int M, const int2* __restrict__ chunk, const int* __restrict__ n2e_ptr,
    const int* __restrict__ n2e, const double* __restrict__ fe, const unsigned char* __restrict__ fix,
    const double* __restrict__ u, const double* __restrict__ f, double* __restrict__ y, double* partials, int mode,
    const int* flags, int guard) {
  if (guard && flags[0]) return;
  __shared__ double sv[kGatherWaves][kChunkInc][3];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  // XCD-contiguous chunk ranges: neighbouring chunks read the same element
  // contribution lines, which then meet in one L2
  const int nb = (n_chunks + kGatherWaves - 1) / kGatherWaves;
  const int lb = slice_of_block(nb, 1);
  const int c = lb * kGatherWaves + w;
  const bool live = lb < nb && c < n_chunks;
  // chunk c: nodes [a.x, b.x), incidences [a.y, b.y).  Every load below is
  // unconditional with a clamped index (node 0 and incidence 0 always exist)
  // and its value masked afterwards: the three dependent rounds (chunk ->
  // node operands and incidence codes -> element values) are then each
  // issued together, where loads under per-lane branches were waited on one
  // by one (eight serial latencies per wavefront).
  const int cc0 = live ? c : 0;
  const int2 a0 = chunk[cc0], b0 = chunk[cc0 + 1];
  const int2 a = live ? a0 : make_int2(0, 0);
  const int2 b = live ? b0 : make_int2(0, 0);
  const int k0 = a.y, k1 = b.y;
  const bool big = k1 - k0 > kChunkInc;  // then the chunk is one node
  // this lane's node and its operands, loaded before the incidences
  const int n = a.x + l;
  const bool has = n < b.x;
  const int nn = has ? n : 0;
  const int p0 = n2e_ptr[nn], p1 = n2e_ptr[nn + 1];
  const int mraw = fix ? fix[nn] : 0;
  double ur[3], fr[3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    ur[cc] = u[(size_t)nn * 3 + cc];
    fr[cc] = mode == 2 ? f[(size_t)nn * 3 + cc] : 0.0;
  }
  int code[kIncPerLane];
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const int k = k0 + l + 64 * q;
    code[q] = n2e[(!big && k < k1) ? k : 0];  // e * 10 + local
  }
  double v[kIncPerLane][3];
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const double* p = fe + ((size_t)(code[q] % 10) * M + code[q] / 10) * 3;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) v[q][cc] = p[cc];
  }
  const int j0 = has ? p0 - k0 : 0, j1 = has ? p1 - k0 : 0, m = has ? mraw : 0;
  double uu[3] = {0.0, 0.0, 0.0}, ff[3] = {0.0, 0.0, 0.0};
  if (has) {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      if (mode == 1 || ((m >> cc) & 1)) uu[cc] = ur[cc];
      if (mode == 2) ff[cc] = fr[cc];
    }
  }
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const int k = k0 + l + 64 * q;
    const bool ok = !big && k < k1;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) sv[w][l + 64 * q][cc] = ok ? v[q][cc] : 0.0;
  }
  __syncthreads();
  double dot = 0.0;
  if (has) {
    double s[3] = {0.0, 0.0, 0.0};
    if (big) {
      for (int j = k0; j < k1; ++j) {
        const int code = n2e[j];
        const double* p = fe + ((size_t)(code % 10) * M + code / 10) * 3;
        s[0] += p[0];
        s[1] += p[1];
        s[2] += p[2];
      }
    } else {
      for (int j = j0; j < j1; ++j) {
        s[0] += sv[w][j][0];
        s[1] += sv[w][j][1];
        s[2] += sv[w][j][2];
      }
    }
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const size_t d = (size_t)n * 3 + cc;
      const double ku = (m >> cc) & 1 ? uu[cc] : s[cc];
      if (mode == 2) {
        const double r = ff[cc] - ku;
        y[d] = r;
        dot += r * r;
      } else {
        y[d] = ku;
        if (mode == 1) dot += uu[cc] * ku;
      }
    }
  }
  if (mode != 0) {
    __shared__ double sm[kGatherWaves];
    double v[1] = {block_sum(dot, sm)};
    store_block_partials<1>(v, partials);
  }
}

// The same gather for the V-cycle's finest level, on the scaled operator
// A = diag(sc) K diag(sc): for DOF k of node n, a_k = sc_k (K (sc .* x))_k
// ((K u)_k = u_k for a fixed DOF), then the epilogue Op (a Chebyshev step or
// the defect, as in the sliced-ELL products) consumes a_k with its operands
// loaded before the incidences.  Output order is node order, so every
// operand and result stream is contiguous.
template <typename T, typename Op>
__global__ __launch_bounds__(64 * kGatherWaves) void mf_gather_op_kernel(
    int n_chunks, int M, const int2* __restrict__ chunk, const int* __restrict__ n2e_ptr,
    const int* __restrict__ n2e, const T* __restrict__ fe, const unsigned char* __restrict__ fix,
    const T* __restrict__ sc, const T* __restrict__ x, Op op, const int* skip) {
  if (skip && *skip) return;
  __shared__ T sv[kGatherWaves][kChunkInc][3];
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  // XCD-contiguous chunk ranges: neighbouring chunks read the same element
  // contribution lines, which then meet in one L2
  const int nb = (n_chunks + kGatherWaves - 1) / kGatherWaves;
  const int lb = slice_of_block(nb, 1);
  const int c = lb * kGatherWaves + w;
  const bool live = lb < nb && c < n_chunks;
  // branch-free loads with clamped indices, as in mf_gather_kernel
  const int cc0 = live ? c : 0;
  const int2 a0 = chunk[cc0], b0 = chunk[cc0 + 1];
  const int2 a = live ? a0 : make_int2(0, 0);
  const int2 b = live ? b0 : make_int2(0, 0);
  const int k0 = a.y, k1 = b.y;
  const bool big = k1 - k0 > kChunkInc;
  const int n = a.x + l;
  const bool has = n < b.x;
  const int nn = has ? n : 0;
  const int p0 = n2e_ptr[nn], p1 = n2e_ptr[nn + 1];
  const int mraw = fix ? fix[nn] : 0;
  T s3[3] = {T(1), T(1), T(1)}, xr[3];
  typename Op::Pre pre[3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const long long k = (long long)nn * 3 + cc;
    op.load(k, pre[cc]);
    if (sc) s3[cc] = sc[k];
    xr[cc] = x[k];
  }
  int code[kIncPerLane];
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const int k = k0 + l + 64 * q;
    code[q] = n2e[(!big && k < k1) ? k : 0];  // e * 10 + local
  }
  T v[kIncPerLane][3];
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const T* p = fe + ((size_t)(code[q] % 10) * M + code[q] / 10) * 3;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) v[q][cc] = p[cc];
  }
  const int j0 = has ? p0 - k0 : 0, j1 = has ? p1 - k0 : 0, m = has ? mraw : 0;
  T xf[3];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) xf[cc] = (m >> cc) & 1 ? xr[cc] : T(0);
#pragma unroll
  for (int q = 0; q < kIncPerLane; ++q) {
    const int k = k0 + l + 64 * q;
    const bool ok = !big && k < k1;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) sv[w][l + 64 * q][cc] = ok ? v[q][cc] : T(0);
  }
  __syncthreads();
  if (!has) return;
  T acc[3] = {T(0), T(0), T(0)};
  if (big) {
    for (int j = k0; j < k1; ++j) {
      const int code = n2e[j];
      const T* p = fe + ((size_t)(code % 10) * M + code / 10) * 3;
      acc[0] += p[0];
      acc[1] += p[1];
      acc[2] += p[2];
    }
  } else {
    for (int j = j0; j < j1; ++j) {
      acc[0] += sv[w][j][0];
      acc[1] += sv[w][j][1];
      acc[2] += sv[w][j][2];
    }
  }
  double dots[2] = {0.0, 0.0};
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const T ku = (m >> cc) & 1 ? s3[cc] * xf[cc] : acc[cc];
    op.out((long long)n * 3 + cc, s3[cc] * ku, dots, pre[cc]);
  }
}

// ---------------------------------------------------------------- patch layout
// One workgroup of 4 P lanes per patch of P consecutive element slots
// (elements sorted along a Morton curve, so a patch is a compact piece of the
// mesh), one quad of lanes per element (lane g: Gauss point g):
//  1. loads addressed by the patch index alone are issued first: each
//     element's 16-bit local connectivity and quadrature data, the patch's
//     slot codes (into LDS) and its local nodes' global ids and list starts
//     (fixed stride per patch); then those addressed through them: the
//     scaled nodal values sc .* U of the local nodes (into LDS, four entries
//     per node; up to three nodes per lane, loads issued together) and,
//     unfolded, the element moduli;
//  2. each quad computes its element's forces as c3d10_pa_elem_kernel does,
//     every lane reading the element's ten nodal values from LDS, and keeps
//     them in LDS;
//  3. each local node adds its element forces in slot order (slot codes
//     ascending, from LDS) and writes the sum to its partial slot
//     p stride + j (contiguous per patch).
// The quadrature data is one record per slot: Ji and det J / 24 ([M][10],
// the moduli read through the slot's element id), or Folded, Ji and the
// weighted moduli w c11, w c12, w c44 ([M][12], made when a smoother is
// attached).
// The element forces never leave the chip; a node interior to a patch has a
// single partial, a node on a patch boundary one per patch that touches it.
constexpr int kStageSlots = 3;  // local nodes per lane in step 1: 3 * 4 P >= the stride

__device__ __forceinline__ void lds_node(const float* xs, int li, float (&u)[3]) {
  const float4 v = *reinterpret_cast<const float4*>(xs + 4 * li);
  u[0] = v.x;
  u[1] = v.y;
  u[2] = v.z;
}
__device__ __forceinline__ void lds_node(const double* xs, int li, double (&u)[3]) {
  const double2 v = *reinterpret_cast<const double2*>(xs + 4 * li);
  u[0] = v.x;
  u[1] = v.y;
  u[2] = xs[4 * li + 2];
}
__device__ __forceinline__ void lds_store_node(float* xs, int j, const float (&v)[3]) {
  *reinterpret_cast<float4*>(xs + 4 * j) = make_float4(v[0], v[1], v[2], 0.0f);
}
__device__ __forceinline__ void lds_store_node(double* xs, int j, const double (&v)[3]) {
  *reinterpret_cast<double2*>(xs + 4 * j) = make_double2(v[0], v[1]);
  xs[4 * j + 2] = v[2];
}

template <typename T, int P, bool Folded>
__global__ __launch_bounds__(4 * P) void mf_patch_kernel(
    int n_patch, int M, int stride, const int* __restrict__ pnode, const unsigned short* __restrict__ pconn,
    const unsigned short* __restrict__ poff, const unsigned short* __restrict__ pslot, const int* __restrict__ pe,
    const T* __restrict__ qdp, const double* __restrict__ E, const double* __restrict__ nu,
    const T* __restrict__ sc, const T* __restrict__ U, T* __restrict__ part, const int* skip) {
  constexpr int NT = 4 * P;
  if (skip && *skip) return;
  extern __shared__ __align__(16) unsigned char patch_lds[];
  T* xs = reinterpret_cast<T*>(patch_lds);                              // [stride][4] scaled nodal values
  T* fl = xs + 4 * stride;                                               // [P][10][3] element forces
  unsigned short* ls = reinterpret_cast<unsigned short*>(fl + 30 * P);  // [10 P] slot codes
  unsigned short* lo = ls + 10 * P;                                       // [stride] list starts
  const int p = slice_of_block(n_patch, 1);  // XCD-contiguous patch ranges
  if (p >= n_patch) return;
  const int t = (int)threadIdx.x;
  // Lane offsets are unsigned 32-bit from uniform per-array bases (the
  // set-up keeps 12 M and 3 n_patch stride below 2^31), so the loads take
  // the scalar-base + vector-offset form and share their offsets.
  const int e0 = p * P, ne = min(P, M - e0);
  const size_t pb = (size_t)p * stride;
  // step 1, first round: addressed by the patch index
  const int el = t >> 2, g = t & 3;
  const bool live = el < ne;
  const unsigned su = (unsigned)(e0 + (live ? el : 0));  // clamped: the loads stay unconditional
  int lc[10];
  {
    const unsigned short* pc = pconn + (size_t)su * 10;
#pragma unroll
    for (int i = 0; i < 10; ++i) lc[i] = pc[i];
  }
  // the slot's quadrature record: 16-byte aligned (48 B folded FP32, 80 B FP64)
  constexpr int QW = Folded ? 12 : 10;
  T qv[QW];
  {
    const T* q = qdp + (size_t)su * QW;
#pragma unroll
    for (int k = 0; k < QW; ++k) qv[k] = q[k];
  }
  T Ji[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Ji[k] = qv[k];
  T w11 = T(0), w12 = T(0), w44 = T(0), wd = T(0);
  int ee = 0;
  if constexpr (Folded) {
    w11 = qv[QW - 3];
    w12 = qv[QW - 2];
    w44 = qv[QW - 1];
  } else {
    wd = qv[9];
    ee = pe[su];
  }
  const unsigned short* psl = pslot + (size_t)e0 * 10;
  unsigned short code[3];  // slot codes t, t + NT, t + 2 NT (10 P < 3 NT)
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const unsigned q = (unsigned)(t + k * NT);
    code[k] = psl[q < (unsigned)(10 * ne) ? q : 0u];
  }
  const int* pn = pnode + pb;
  const unsigned short* po = poff + pb;
  int gn[kStageSlots];
  unsigned short off[kStageSlots];
#pragma unroll
  for (int k = 0; k < kStageSlots; ++k) {
    const unsigned j = (unsigned)min(t + k * NT, stride - 1);
    gn[k] = pn[j];
    off[k] = po[j];
  }
  // second round: addressed through the first
  double Ev = 0.0, nv = 0.0;
  if constexpr (!Folded) {
    Ev = E[ee];
    nv = nu[ee];
  }
  T xv[kStageSlots][3];
#pragma unroll
  for (int k = 0; k < kStageSlots; ++k) {
    const size_t n3 = 3 * (size_t)max(gn[k], 0);
    const T* uk = U + n3;
    const T* sk = sc + n3;
#pragma unroll
    for (int c = 0; c < 3; ++c) xv[k][c] = uk[c] * sk[c];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (t + k * NT < 10 * ne) ls[t + k * NT] = code[k];
#pragma unroll
  for (int k = 0; k < kStageSlots; ++k) {
    const int j = t + k * NT;
    if (j < stride) {
      lo[j] = off[k];
      lds_store_node(xs, j, xv[k]);
    }
  }
  __syncthreads();
  // step 2: the element forces
  if constexpr (!Folded) weighted_moduli<T>(wd, Ev, nv, w11, w12, w44);
  T c4[4], cc[4];
  gauss_point_consts<T>(g, c4, cc);
  T out[3][3];
  pa_forces<T>(g, c4, cc, Ji, w11, w12, w44, [&](int i, T (&u)[3]) { lds_node(xs, lc[i], u); }, out);
  if (live) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int i = g + 4 * j;
      if (i < 10) {
#pragma unroll
        for (int c = 0; c < 3; ++c) fl[(el * 10 + i) * 3 + c] = out[j][c];
      }
    }
  }
  __syncthreads();
  // step 3: per local node, its element forces in slot order (the entry
  // after the last node holds the end of its list)
#pragma unroll
  for (int k = 0; k < kStageSlots; ++k) {
    const int j = t + k * NT;
    if (j < stride && gn[k] >= 0) {
      const int o0 = lo[j], o1 = lo[j + 1];
      T a0 = T(0), a1 = T(0), a2 = T(0);
      for (int q = o0; q < o1; ++q) {
        const int f = 3 * ls[q];
        a0 += fl[f];
        a1 += fl[f + 1];
        a2 += fl[f + 2];
      }
      T* o = part + pb * 3;
      const unsigned j3 = 3u * (unsigned)j;
      o[j3] = a0;
      o[j3 + 1] = a1;
      o[j3 + 2] = a2;
    }
  }
}

// Folded FP32 quadrature records of the patch layout: per slot the rounded
// inverse Jacobian and w c11, w c12, w c44 (formed in FP64 from the slot's
// element moduli), [M][12].
__global__ void qdata_fold_kernel(int M, const int* __restrict__ pe, const double* __restrict__ qdp,
                                  const double* __restrict__ E, const double* __restrict__ nu,
                                  float* __restrict__ out) {
  for (int sl = blockIdx.x * blockDim.x + threadIdx.x; sl < M; sl += gridDim.x * blockDim.x) {
    const double* q = qdp + (size_t)sl * 10;
    float* o = out + (size_t)sl * 12;
    for (int k = 0; k < 9; ++k) o[k] = (float)q[k];
    const double wd = q[9], Ev = E[pe[sl]], nv = nu[pe[sl]];
    const double f = Ev / ((1.0 + nv) * (1.0 - 2.0 * nv));
    o[9] = (float)(wd * f * (1.0 - nv));
    o[10] = (float)(wd * f * nv);
    o[11] = (float)(wd * Ev / (2.0 * (1.0 + nv)));
  }
}

// Node-ordered completion of the patch layout: node n's partial sums (its
// head: the first three slots and the count, the rest in an overflow list;
// all in patch order) are added, the assembled operator's Dirichlet rule
// applied (a fixed DOF k gives s_k^2 x_k, the unit diagonal of the scaled
// operator) and the epilogue Op consumes a_k = s_k (K (sc .* x))_k as in
// mf_gather_op_kernel (s = 1 without a scaling).  One lane per node; every
// operand and result streams in node order, and the partial sums are one
// dependent load behind the head.  Ops with dot products store one partial
// per block of the padded grid (zero for blocks without nodes).
constexpr int kNodeThreads = 256;

template <typename T, typename Op>
__global__ __launch_bounds__(kNodeThreads) void mf_node_kernel(int n_nodes, const int4* __restrict__ nhead,
                                                              const int* __restrict__ novf, const T* __restrict__ part,
                                                              const unsigned char* __restrict__ fix,
                                                              const T* __restrict__ sc, const T* __restrict__ x,
                                                              Op op, const int* skip) {
  if (skip && *skip) return;
  const int nblk = (n_nodes + kNodeThreads - 1) / kNodeThreads;
  const int lb = slice_of_block(nblk, 1);
  const int n = lb * kNodeThreads + (int)threadIdx.x;
  const bool has = lb < nblk && n < n_nodes;
  const int nn = has ? n : 0;
  const int4 h = nhead[nn];
  const int mraw = fix ? fix[nn] : 0;
  T s3[3] = {T(1), T(1), T(1)}, xr[3];
  typename Op::Pre pre[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long k = (long long)nn * 3 + c;
    op.load(k, pre[c]);
    if (sc) s3[c] = sc[k];
    xr[c] = x[k];
  }
  const int hi[3] = {h.x, h.y, h.z};
  T v[3][3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) v[q][c] = part[(size_t)max(hi[q], 0) * 3 + c];
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (hi[q] >= 0) acc[c] = q == 0 ? v[q][c] : acc[c] + v[q][c];
  if (h.w < 0) {  // more than three partials
    const int k = -h.w - 1, cnt = novf[k];
    for (int q = 1; q <= cnt; ++q) {
      const T* w = part + (size_t)novf[k + q] * 3;
      acc[0] += w[0];
      acc[1] += w[1];
      acc[2] += w[2];
    }
  }
  double dots[2] = {0.0, 0.0};
  if (has) {
    const int m = mraw;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T ku = (m >> c) & 1 ? s3[c] * xr[c] : acc[c];
      op.out((long long)n * 3 + c, s3[c] * ku, dots, pre[c]);
    }
  }
  if constexpr (Op::kDots > 0) {
    if (op.partials) {
      __shared__ double sm[kNodeThreads / 64];
      double vv[Op::kDots];
#pragma unroll
      for (int k = 0; k < Op::kDots; ++k) vv[k] = block_sum(dots[k], sm);
      store_block_partials<Op::kDots>(vv, op.partials);
    }
  }
}

// The modes of mf_apply as an epilogue (mode 0: y = K u; 1: y = K u and the
// dot u.y; 2: y = f - K u and the dot y.y).
struct OpMfMode {
  static constexpr int kDots = 1;
  const double* u;
  const double* f;
  double* y;
  int mode;
  double* partials;  // null: no dot
  struct Pre { double u, f; };
  __device__ __forceinline__ void load(long long k, Pre& p) const {
    p.u = u[k];
    p.f = mode == 2 ? f[k] : 0.0;
  }
  __device__ __forceinline__ void out(long long k, double a, double (&d)[2], const Pre& p) const {
    if (mode == 2) {
      const double r = p.f - a;
      y[k] = r;
      d[0] += r * r;
    } else {
      y[k] = a;
      if (mode == 1) d[0] += p.u * a;
    }
  }
};

// qdp[s][k] = qd[k][pe[s]] (affine quadrature data as one record per slot)
__global__ void qdata_permute_kernel(int M, const int* __restrict__ pe, const double* __restrict__ qd,
                                     double* __restrict__ qdp) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 10LL * M;
       i += (long long)gridDim.x * blockDim.x) {
    const long long sl = i / 10, k = i % 10;
    qdp[i] = qd[k * M + pe[sl]];
  }
}

// out_k = 0 for a DOF fixed in node_fix (bit k % 3 of node k / 3), else
// scale_k (1 without a scale): the element kernel's input scaling.
template <typename T>
__global__ void dof_scale_kernel(long long n, const double* __restrict__ scale, const unsigned char* __restrict__ fix,
                                 T* __restrict__ out) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    const bool fixed = fix && ((fix[k / 3] >> (k % 3)) & 1);
    out[k] = fixed ? T(0) : (scale ? T(scale[k]) : T(1));
  }
}

__global__ void round_to_float_kernel(long long n, const double* __restrict__ src, float* __restrict__ dst) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// ---------------------------------------------------------------- host wrappers
int launch_ke(int M, const int* conn, const double* nodes, const double* E, const double* nu, double* Ke, int* bad,
              hipStream_t s) {
  if (M <= 0) return kOk;
  int r = upload_element_consts(s);
  if (r) return r;
  if (g_ke_variant == 1)
    hipLaunchKernelGGL(c3d10_ke_mfma_kernel, dim3(ceil_div(M, kKeMfElem)), dim3(64 * kKeMfElem), 0, s, M, conn, nodes,
                       E, nu, Ke, bad);
  else
    hipLaunchKernelGGL(c3d10_ke_kernel, dim3(ceil_div(M, kElemPerBlock)), dim3(256), 0, s, M, conn, nodes, E, nu, Ke,
                       bad);
  return static_cast<int>(hipGetLastError());
}

int launch_fint_elem(int M, const int* conn, const double* nodes, const double* E, const double* nu, const double* U,
                     int stride, double* fe, hipStream_t s) {
  if (M <= 0) return kOk;
  int r = upload_element_consts(s);
  if (r) return r;
  hipLaunchKernelGGL(c3d10_fint_elem_kernel, dim3(ceil_div(M, 128)), dim3(128), 0, s, M, conn, nodes, E, nu, U,
                     stride, fe);
  return static_cast<int>(hipGetLastError());
}

int launch_gather_node(int n_nodes, const int* n2e_ptr, const int* n2e, const double* fe, double* out, int stride,
                       int accumulate, hipStream_t s) {
  if (n_nodes <= 0) return kOk;
  hipLaunchKernelGGL(gather_node_kernel, dim3(ceil_div(n_nodes, 256)), dim3(256), 0, s, n_nodes, n2e_ptr, n2e, fe,
                     out, stride, accumulate);
  return static_cast<int>(hipGetLastError());
}

// Gather chunks: consecutive nodes, at most 64 per chunk and at most
// kChunkInc incidences in all, or one node alone when it has more.
std::vector<int> mf_chunks(const std::vector<int>& ptr) {  // first node of each chunk, then n
  const int n = (int)ptr.size() - 1;
  std::vector<int> c{0};
  int start = 0;
  while (start < n) {
    int end = start + 1;
    if (ptr[end] - ptr[start] <= kChunkInc)
      while (end < n && end - start < 64 && ptr[end + 1] - ptr[start] <= kChunkInc) ++end;
    c.push_back(end);
    start = end;
  }
  return c;
}

// ---------------------------------------------------------------- patch layout (host set-up)
static void mf_patch_release(MfOp& op) {
  for (void* p : {(void*)op.pe, (void*)op.pnode, (void*)op.pconn, (void*)op.poff, (void*)op.pslot, (void*)op.nhead,
                  (void*)op.novf, (void*)op.qdp, (void*)op.part})
    if (p) (void)hipFree(p);
  op.pe = op.pnode = op.novf = nullptr;
  op.nhead = nullptr;
  op.pconn = op.poff = op.pslot = nullptr;
  op.qdp = op.part = nullptr;
  op.patch = op.n_patch = op.lmax = op.stride = op.n_local = 0;
  ++op.gen;
}

// LDS of one patch workgroup: staged nodal values and element forces
static size_t patch_lds_bytes(const MfOp& op, size_t elem) {
  return (4 * (size_t)op.stride + 30 * (size_t)op.patch) * elem + 2 * (10 * (size_t)op.patch + (size_t)op.stride);
}
constexpr size_t kPatchLdsMax = 64 * 1024;
constexpr int kDefaultPatch = 64;

int mf_set_patch(MfOp& op, int elems, hipStream_t s) {
  if (elems < 0 || elems > 256 || elems % 64) return kBadArg;
  mf_patch_release(op);
  if (elems == 0) return kOk;
  if (!op.affine || op.M <= 0 || !op.qd || op.order.size() != (size_t)op.M) return kBadMode;
  const int M = op.M, N = op.n_nodes, P = elems;
  std::vector<int> conn((size_t)M * 10);
  WN_HIP(hipMemcpy(conn.data(), op.conn, sizeof(int) * conn.size(), hipMemcpyDeviceToHost));
  const std::vector<int>& order = op.order;
  int lmax = 0;
  const long long nl = mf_patch_sizes(M, N, conn.data(), order.data(), P, &lmax);
  if (nl < 0) return kBadArg;
  op.patch = P;
  op.n_patch = ceil_div(M, P);
  op.lmax = lmax;
  op.stride = (lmax + 1 + 3) / 4 * 4;
  op.n_local = (int)nl;
  // the FP64 patch must fit one workgroup's LDS, and step 1 of the patch
  // kernel stages at most kStageSlots local nodes per lane
  if (patch_lds_bytes(op, sizeof(double)) > kPatchLdsMax || op.stride > kStageSlots * 4 * P ||
      12LL * M >= (1LL << 31) || (long long)op.n_patch * op.stride * 3 >= (1LL << 31)) {
    mf_patch_release(op);
    return kBadSize;
  }
  const size_t cap = (size_t)op.n_patch * op.stride;
  std::vector<int> pnode(cap), nhead((size_t)N * 4), novf((size_t)nl + 1);
  std::vector<unsigned short> pconn((size_t)M * 10), pslot((size_t)M * 10), poff(cap);
  const long long nov = mf_patches(M, N, conn.data(), order.data(), P, op.stride, pnode.data(), pconn.data(),
                                   poff.data(), pslot.data(), nhead.data(), novf.data());
  if (nov < 0) {
    mf_patch_release(op);
    return kBadArg;
  }
  novf.resize((size_t)std::max<long long>(nov, 1));
  auto up = [&](auto*& dst, const auto& v) -> int {
    using E = typename std::decay_t<decltype(v)>::value_type;
    WN_HIP(hipMalloc(&dst, sizeof(E) * std::max<size_t>(v.size(), 1)));
    WN_HIP(hipMemcpy(dst, v.data(), sizeof(E) * v.size(), hipMemcpyHostToDevice));
    return kOk;
  };
  int r;
  if ((r = up(op.pe, order)) || (r = up(op.pnode, pnode)) || (r = up(op.pconn, pconn)) || (r = up(op.poff, poff)) ||
      (r = up(op.pslot, pslot)) || (r = up(op.nhead, nhead)) || (r = up(op.novf, novf))) {
    mf_patch_release(op);
    return r;
  }
  hipError_t e = hipMalloc(&op.qdp, sizeof(double) * 10 * (size_t)M);
  if (e == hipSuccess) e = hipMalloc(&op.part, sizeof(double) * 3 * cap);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(qdata_permute_kernel, dim3(vec_grid(10LL * M)), dim3(256), 0, s, M, op.pe, op.qd, op.qdp);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    mf_patch_release(op);
    return static_cast<int>(e);
  }
  return kOk;
}

int mf_setup(MfOp& op, const double* nodes, int* bad, hipStream_t s) {
  mf_release(op);
  if (op.n_nodes <= 0) return kBadArg;
  std::vector<int> ptr((size_t)op.n_nodes + 1);
  WN_HIP(hipMemcpy(ptr.data(), op.n2e_ptr, sizeof(int) * ptr.size(), hipMemcpyDeviceToHost));
  if (ptr[0] != 0 || ptr[op.n_nodes] != 10LL * op.M) return kBadArg;
  for (int i = 0; i < op.n_nodes; ++i)
    if (ptr[i + 1] < ptr[i]) return kBadArg;
  const std::vector<int> c = mf_chunks(ptr);
  op.n_chunks = (int)c.size() - 1;
  std::vector<int2> c2(c.size());
  for (size_t i = 0; i < c.size(); ++i) c2[i] = make_int2(c[i], ptr[(size_t)c[i]]);
  WN_HIP(hipMalloc(&op.chunk, sizeof(int2) * c2.size()));
  WN_HIP(hipMemcpy(op.chunk, c2.data(), sizeof(int2) * c2.size(), hipMemcpyHostToDevice));
  if (op.M <= 0) return kOk;
  int r = upload_element_consts(s);
  if (r) return r;
  const long long n = 3LL * op.n_nodes;
  WN_HIP(hipMalloc(&op.qd, sizeof(double) * 40 * (size_t)op.M));
  WN_HIP(hipMalloc(&op.fe, sizeof(double) * 30 * (size_t)op.M));
  WN_HIP(hipMalloc(&op.mask, sizeof(double) * n));
  int* curved = nullptr;
  WN_HIP(hipMalloc(&curved, sizeof(int)));
  if (const hipError_t me = hipMemsetAsync(curved, 0, sizeof(int), s); me != hipSuccess) {
    (void)hipFree(curved);
    return static_cast<int>(me);
  }
  hipLaunchKernelGGL(c3d10_qdata_kernel, dim3(ceil_div(op.M, 128)), dim3(128), 0, s, op.M, op.conn, nodes, op.qd, bad,
                     curved);
  hipLaunchKernelGGL(dof_scale_kernel<double>, dim3(vec_grid(n)), dim3(256), 0, s, n, nullptr, op.fix, op.mask);
  int h_curved = 1;
  const hipError_t ce = hipMemcpyAsync(&h_curved, curved, sizeof(int), hipMemcpyDeviceToHost, s);
  const hipError_t se = ce == hipSuccess ? hipStreamSynchronize(s) : ce;
  (void)hipFree(curved);
  WN_HIP(se);
  if (!h_curved) {  // affine mesh: keep one set of quadrature data per element
    double* qa = nullptr;
    WN_HIP(hipMalloc(&qa, sizeof(double) * 10 * (size_t)op.M));
    hipLaunchKernelGGL(qdata_affine_kernel, dim3(vec_grid(10LL * op.M)), dim3(256), 0, s, 10LL * op.M, op.qd, qa);
    WN_HIP(hipStreamSynchronize(s));
    (void)hipFree(op.qd);
    op.qd = qa;
    op.affine = 1;
  }
  WN_HIP(hipGetLastError());
  if (op.affine) {  // the patch layout, when one patch's FP64 workgroup fits the LDS
    std::vector<int> conn((size_t)op.M * 10);
    std::vector<double> X((size_t)op.n_nodes * 3);
    WN_HIP(hipMemcpy(conn.data(), op.conn, sizeof(int) * conn.size(), hipMemcpyDeviceToHost));
    WN_HIP(hipMemcpy(X.data(), nodes, sizeof(double) * X.size(), hipMemcpyDeviceToHost));
    for (int v : conn)
      if (v < 0 || v >= op.n_nodes) return kBadArg;
    op.order.resize((size_t)op.M);
    morton_order(op.M, conn.data(), X.data(), op.order.data());
    r = mf_set_patch(op, kDefaultPatch, s);
    if (r != kOk && r != kBadSize) return r;
  }
  return kOk;
}

void mf_release(MfOp& op) {
  mf_patch_release(op);
  op.order.clear();
  for (void* p : {(void*)op.qd, (void*)op.fe, (void*)op.mask, (void*)op.chunk})
    if (p) (void)hipFree(p);
  op.qd = op.fe = op.mask = nullptr;
  op.affine = 0;
  op.chunk = nullptr;
  op.n_chunks = 0;
}

// blocks of the node-ordered kernel that completes the patch layout
static int node_grid(int n_nodes) { return 8 * ceil_div(std::max(1, ceil_div(n_nodes, kNodeThreads)), 8); }
static int patch_grid(int n_patch) { return 8 * ceil_div(std::max(1, n_patch), 8); }

// the patch kernel for the layout's patch size
template <typename T, bool Folded>
static void launch_patch(const MfOp& op, const T* qdp, const T* sc, const T* U, T* part, const int* skip,
                         hipStream_t s) {
  const dim3 grid(patch_grid(op.n_patch));
  const size_t lds = patch_lds_bytes(op, sizeof(T));
#define WN_PATCH_LAUNCH(PP)                                                                                         \
  hipLaunchKernelGGL((mf_patch_kernel<T, PP, Folded>), grid, dim3(4 * (PP)), lds, s, op.n_patch, op.M, op.stride,   \
                     op.pnode, op.pconn, op.poff, op.pslot, op.pe, qdp, op.E, op.nu, sc, U, part, skip)
  switch (op.patch) {
    case 64: WN_PATCH_LAUNCH(64); break;
    case 128: WN_PATCH_LAUNCH(128); break;
    case 192: WN_PATCH_LAUNCH(192); break;
    default: WN_PATCH_LAUNCH(256); break;
  }
#undef WN_PATCH_LAUNCH
}

// blocks of the per-element path's node gather, padded to a multiple of the 8 XCDs
static int gather_grid(const MfOp& op) { return 8 * ceil_div(std::max(1, ceil_div(op.n_chunks, kGatherWaves)), 8); }

// blocks of the final kernel of mf_apply (the node kernel of the patch
// layout, or the gather)
int mf_grid(const MfOp& op) { return op.patch ? node_grid(op.n_nodes) : gather_grid(op); }
static int elem_grid(int M) { return 8 * ceil_div(ceil_div(4L * M, 256), 8); }

int mf_apply(const MfOp& op, const double* u, const double* f, double* y, double* partials, int mode,
             const int* flags, int guard, hipStream_t s) {
  if (!op.chunk) return kNotReady;
  if (op.patch) {
    if (!op.mask) return kNotReady;
    const int* skip = guard ? flags : nullptr;
    launch_patch<double, false>(op, op.qdp, op.mask, u, op.part, skip, s);
    const OpMfMode o{u, f, y, mode, mode ? partials : nullptr};
    hipLaunchKernelGGL((mf_node_kernel<double, OpMfMode>), dim3(node_grid(op.n_nodes)), dim3(kNodeThreads), 0, s,
                       op.n_nodes, op.nhead, op.novf, op.part, op.fix, nullptr, u, o, skip);
    return static_cast<int>(hipGetLastError());
  }
  if (op.M > 0) {
    if (!op.qd || !op.fe || !op.mask) return kNotReady;
    if (op.affine)
      hipLaunchKernelGGL((c3d10_pa_elem_kernel<double, true>), dim3(elem_grid(op.M)), dim3(256), 0, s, op.M,
                         op.conn, op.qd, op.E, op.nu, op.mask, u, op.fe, flags, guard);
    else
      hipLaunchKernelGGL((c3d10_pa_elem_kernel<double, false>), dim3(elem_grid(op.M)), dim3(256), 0, s, op.M,
                         op.conn, op.qd, op.E, op.nu, op.mask, u, op.fe, flags, guard);
  }
  hipLaunchKernelGGL(mf_gather_kernel, dim3(gather_grid(op)), dim3(64 * kGatherWaves), 0, s, op.n_chunks, op.M, op.chunk,
                     op.n2e_ptr, op.n2e, op.fe, op.fix, u, f, y, partials, mode, flags, guard);
  return static_cast<int>(hipGetLastError());
}

int launch_stress(int M, const int* conn, const double* nodes, const double* E, const double* nu, const double* U,
                  double* gp_sig, double* gp_vm, double* cell_sig, double* cell_vm, int* bad, hipStream_t s) {
  if (M <= 0) return kOk;
  int r = upload_element_consts(s);
  if (r) return r;
  hipLaunchKernelGGL(c3d10_stress_kernel, dim3(ceil_div(M, 128)), dim3(128), 0, s, M, conn, nodes, E, nu, U, gp_sig,
                     gp_vm, cell_sig, cell_vm, bad);
  return static_cast<int>(hipGetLastError());
}

int launch_nodal_average(int n_nodes, const int* n2e_ptr, const int* n2e, const double* cell_vm, double* nodal_vm,
                         hipStream_t s) {
  if (n_nodes <= 0) return kOk;
  hipLaunchKernelGGL(nodal_average_kernel, dim3(ceil_div(n_nodes, 256)), dim3(256), 0, s, n_nodes, n2e_ptr, n2e,
                     cell_vm, nodal_vm);
  return static_cast<int>(hipGetLastError());
}

int launch_adjoint_dot(int M, const int* conn, const double* E, const double* fe, const double* lam, double* grad,
                       hipStream_t s) {
  if (M <= 0) return kOk;
  hipLaunchKernelGGL(adjoint_dot_kernel, dim3(ceil_div(M, 256)), dim3(256), 0, s, M, conn, E, fe, lam, grad);
  return static_cast<int>(hipGetLastError());
}

// ---------------------------------------------------------------- V-cycle smoothing on the matrix-free operator
void mf_smooth_release(MfSmooth& m) {
  for (void* p : {(void*)m.s64e, (void*)m.s32, (void*)m.s32e, (void*)m.qd32, (void*)m.fe32, (void*)m.qdp32,
                  (void*)m.part32})
    if (p) (void)hipFree(p);
  m = MfSmooth{};
}

int mf_smooth_attach(MfSmooth& m, const MfOp* op, const double* scale, int want32, hipStream_t s) {
  mf_smooth_release(m);
  if (!op) return kOk;  // detach
  if (!op->chunk || (op->M > 0 && (!op->qd || !op->fe)) || !scale) return kBadArg;
  const long long n = 3LL * op->n_nodes;
  const size_t nv = (size_t)std::max<long long>(n, 1);
  WN_HIP(hipMalloc(&m.s64e, sizeof(double) * nv));
  hipLaunchKernelGGL(dof_scale_kernel<double>, dim3(vec_grid(n)), dim3(256), 0, s, n, scale, op->fix, m.s64e);
  if (want32) {
    WN_HIP(hipMalloc(&m.s32, sizeof(float) * nv));
    WN_HIP(hipMalloc(&m.s32e, sizeof(float) * nv));
    const long long nq = (op->affine ? 10LL : 40LL) * op->M;
    WN_HIP(hipMalloc(&m.qd32, sizeof(float) * std::max<long long>(nq, 1)));
    WN_HIP(hipMalloc(&m.fe32, sizeof(float) * 30 * std::max<size_t>((size_t)op->M, 1)));
    hipLaunchKernelGGL(round_to_float_kernel, dim3(vec_grid(n)), dim3(256), 0, s, n, scale, m.s32);
    hipLaunchKernelGGL(dof_scale_kernel<float>, dim3(vec_grid(n)), dim3(256), 0, s, n, scale, op->fix, m.s32e);
    if (op->M > 0)
      hipLaunchKernelGGL(round_to_float_kernel, dim3(vec_grid(nq)), dim3(256), 0, s, nq, op->qd, m.qd32);
    if (op->patch) {
      WN_HIP(hipMalloc(&m.qdp32, sizeof(float) * 12 * (size_t)op->M));
      WN_HIP(hipMalloc(&m.part32, sizeof(float) * 3 * (size_t)op->n_patch * op->stride));
      hipLaunchKernelGGL(qdata_fold_kernel, dim3(vec_grid(op->M)), dim3(256), 0, s, op->M, op->pe, op->qdp, op->E,
                         op->nu, m.qdp32);
    }
  }
  m.gen = op->gen;
  WN_HIP(hipGetLastError());
  m.op = op;
  m.s64 = scale;
  return kOk;
}

template <> bool mf_smooth_ready<double>(const MfSmooth& m) { return m.op && m.s64 && m.s64e; }
template <> bool mf_smooth_ready<float>(const MfSmooth& m) { return m.op && m.s32 && m.s32e && m.qd32 && m.fe32; }

// scaling of the gathered result (the true s) and of the element input (s with fixed DOFs zeroed)
template <typename T> static const T* mf_scale(const MfSmooth& m);
template <> const double* mf_scale<double>(const MfSmooth& m) { return m.s64; }
template <> const float* mf_scale<float>(const MfSmooth& m) { return m.s32; }
template <typename T> static const T* mf_scale_in(const MfSmooth& m);
template <> const double* mf_scale_in<double>(const MfSmooth& m) { return m.s64e; }
template <> const float* mf_scale_in<float>(const MfSmooth& m) { return m.s32e; }
template <typename T> static const T* mf_qd(const MfSmooth& m);
template <> const double* mf_qd<double>(const MfSmooth& m) { return m.op->qd; }
template <> const float* mf_qd<float>(const MfSmooth& m) { return m.qd32; }
template <typename T> static T* mf_fe(const MfSmooth& m);
template <> double* mf_fe<double>(const MfSmooth& m) { return m.op->fe; }
template <> float* mf_fe<float>(const MfSmooth& m) { return m.fe32; }

// the patch layout's FP64 buffers are the operator's; FP32 uses the
// attachment's copies, made for the operator's layout at attach time
template <typename T> static const T* mf_qdp(const MfSmooth& m);
template <> const double* mf_qdp<double>(const MfSmooth& m) { return m.op->patch ? m.op->qdp : nullptr; }
template <> const float* mf_qdp<float>(const MfSmooth& m) {
  return m.op->patch && m.gen == m.op->gen ? m.qdp32 : nullptr;
}
template <typename T> static T* mf_part(const MfSmooth& m);
template <> double* mf_part<double>(const MfSmooth& m) { return m.op->part; }
template <> float* mf_part<float>(const MfSmooth& m) { return m.part32; }

template <typename T, typename Op>
static int mf_smooth_run(const MfSmooth& m, const T* x, const Op& o, const int* skip, hipStream_t s) {
  if (!mf_smooth_ready<T>(m)) return kNotReady;
  const MfOp& op = *m.op;
  const T* sc = mf_scale<T>(m);
  if (const T* qdp = mf_qdp<T>(m)) {
    T* part = mf_part<T>(m);
    launch_patch<T, std::is_same<T, float>::value>(op, qdp, mf_scale_in<T>(m), x, part, skip, s);
    hipLaunchKernelGGL((mf_node_kernel<T, Op>), dim3(node_grid(op.n_nodes)), dim3(kNodeThreads), 0, s, op.n_nodes,
                       op.nhead, op.novf, part, op.fix, sc, x, o, skip);
    return static_cast<int>(hipGetLastError());
  }
  if (op.M > 0 && op.affine)
    hipLaunchKernelGGL((c3d10_pa_elem_kernel<T, true>), dim3(elem_grid(op.M)), dim3(256), 0, s, op.M,
                       op.conn, mf_qd<T>(m), op.E, op.nu, mf_scale_in<T>(m), x, mf_fe<T>(m), skip, skip ? 1 : 0);
  else if (op.M > 0)
    hipLaunchKernelGGL((c3d10_pa_elem_kernel<T, false>), dim3(elem_grid(op.M)), dim3(256), 0, s, op.M,
                       op.conn, mf_qd<T>(m), op.E, op.nu, mf_scale_in<T>(m), x, mf_fe<T>(m), skip, skip ? 1 : 0);
  hipLaunchKernelGGL((mf_gather_op_kernel<T, Op>), dim3(gather_grid(op)), dim3(64 * kGatherWaves), 0, s, op.n_chunks,
                     op.M, op.chunk, op.n2e_ptr, op.n2e, mf_fe<T>(m), op.fix, sc, x, o, skip);
  return static_cast<int>(hipGetLastError());
}

template <typename T>
int mf_smooth_cheb(const MfSmooth& m, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, T c1, T c2,
                   const int* skip, hipStream_t s) {
  if (!dinv) {
    const OpCheb<3, T, true> o{b, nullptr, x_in, x_out, d, c1, c2};
    return mf_smooth_run<T>(m, x_in, o, skip, s);
  }
  const OpCheb<3, T> o{b, dinv, x_in, x_out, d, c1, c2};
  return mf_smooth_run<T>(m, x_in, o, skip, s);
}

template <typename T>
int mf_smooth_defect(const MfSmooth& m, const T* x, const T* b, T* res, const int* skip, hipStream_t s) {
  OpDefect<3, T> o{b, res};
  return mf_smooth_run<T>(m, x, o, skip, s);
}

// A Chebyshev step (as OpCheb) that also forms the partials of
// rho = r . (S z), z = the step's x_out as stored: the V-cycle's last step
// hands the CG iteration its rho without a pass over r, S and z of its own.
// Per DOF the same product as the separate rho kernel, r (S z).
template <typename T, bool Unit = false>
struct OpChebRho {
  static constexpr int kDots = 1;
  OpCheb<3, T, Unit> c;
  const double* r;
  const double* scale;
  double* partials;
  struct Pre {
    typename OpCheb<3, T, Unit>::Pre c;
    double r, s;
  };
  __device__ __forceinline__ void load(long long k, Pre& p) const {
    c.load(k, p.c);
    p.r = r[k];
    p.s = scale ? scale[k] : 1.0;
  }
  __device__ __forceinline__ void out(long long k, T a, double (&dots)[2], const Pre& p) const {
    const T dn = c.c1 * p.c.d + c.c2 * p.c.dinv * (p.c.b - a);
    const T xo = p.c.x + dn;
    c.d[k] = dn;
    c.x_out[k] = xo;
    dots[0] += p.r * (double(xo) * p.s);
  }
};

template <typename T>
int mf_smooth_rho_blocks(const MfSmooth& m) {
  return mf_smooth_ready<T>(m) && mf_qdp<T>(m) ? node_grid(m.op->n_nodes) : 0;
}

template <typename T>
int mf_smooth_cheb_rho(const MfSmooth& m, const T* x_in, T* x_out, const T* b, const T* dinv, T* d, T c1, T c2,
                       const RhoReq& rq, const int* skip, hipStream_t s) {
  const T* qdp = mf_smooth_ready<T>(m) ? mf_qdp<T>(m) : nullptr;
  if (!qdp || !rq.r || !rq.partials) return kBadMode;
  const MfOp& op = *m.op;
  T* part = mf_part<T>(m);
  launch_patch<T, std::is_same<T, float>::value>(op, qdp, mf_scale_in<T>(m), x_in, part, skip, s);
  const dim3 grid(node_grid(op.n_nodes));
  if (!dinv) {
    const OpChebRho<T, true> o{{b, nullptr, x_in, x_out, d, c1, c2}, rq.r, rq.scale, rq.partials};
    hipLaunchKernelGGL((mf_node_kernel<T, OpChebRho<T, true>>), grid, dim3(kNodeThreads), 0, s, op.n_nodes,
                       op.nhead, op.novf, part, op.fix, mf_scale<T>(m), x_in, o, skip);
  } else {
    const OpChebRho<T> o{{b, dinv, x_in, x_out, d, c1, c2}, rq.r, rq.scale, rq.partials};
    hipLaunchKernelGGL((mf_node_kernel<T, OpChebRho<T>>), grid, dim3(kNodeThreads), 0, s, op.n_nodes, op.nhead,
                       op.novf, part, op.fix, mf_scale<T>(m), x_in, o, skip);
  }
  return static_cast<int>(hipGetLastError());
}

template int mf_smooth_cheb<double>(const MfSmooth&, const double*, double*, const double*, const double*, double*,
                                    double, double, const int*, hipStream_t);
template int mf_smooth_cheb<float>(const MfSmooth&, const float*, float*, const float*, const float*, float*, float,
                                   float, const int*, hipStream_t);
template int mf_smooth_defect<double>(const MfSmooth&, const double*, const double*, double*, const int*,
                                      hipStream_t);
template int mf_smooth_defect<float>(const MfSmooth&, const float*, const float*, float*, const int*, hipStream_t);
template int mf_smooth_rho_blocks<double>(const MfSmooth&);
template int mf_smooth_rho_blocks<float>(const MfSmooth&);
template int mf_smooth_cheb_rho<double>(const MfSmooth&, const double*, double*, const double*, const double*,
                                        double*, double, double, const RhoReq&, const int*, hipStream_t);
template int mf_smooth_cheb_rho<float>(const MfSmooth&, const float*, float*, const float*, const float*, float*,
                                       float, float, const RhoReq&, const int*, hipStream_t);

}  // namespace wn
