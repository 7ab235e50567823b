// Co-rotational 3D beam kernels for gfx950 (FP64): internal forces, the
// consistent element tangent (the exact derivative of those forces, by
// forward-mode differentiation with dual numbers; central differences on
// request), and the linear 12x12 beam stiffness.  6 DOF per node
// [ux uy uz rx ry rz].
//
// Internal forces follow the vectorised co-rotational formulation of
// wnfea/solver/fast_kernels.py:451-537 (the path the reference's solver runs):
//   reference triad from the undeformed chord e1 and ref = e_z (|e1_z| < 0.9)
//   or e_y: e2 = unit(ref x e1), e3 = unit(e1 x e2) (fast_kernels.py:297-305);
//   current chord r1; (r2, r3) = the Rodrigues rotation of (e2, e3) taking e1
//   to r1, twisted by the mean axial nodal rotation, re-orthonormalised;
//   elongation u_l = (2 v0.du + du.du)/(L + L0); local rotations
//   th_y = r2.theta + r3.du/L, th_z = r3.theta - r2.du/L, th_x = r1.(theta2 -
//   theta1); q = (N, My1, Mz1, Tx, My2, Mz2) of the linear beam; f = B^T q.
// One thread per element (the element is ~300 flops with transcendental
// functions; no reuse between elements), forces per element in fe[M][12],
// then a deterministic per-node gather (no atomics).
//
// The linear stiffness follows build_element_stiffness_3d_beam and
// build_transformation_matrix (wnfea/solver/assembler.py:13-148), including
// its axis-aligned frame rules.
#include <type_traits>
#include "common.hpp"

namespace wn {

using ::cos;
using ::sin;
using ::sqrt;

// Forward-mode dual number: a value and its derivative along one direction.
struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double a) : v(a), d(0.0) {}
  __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
};
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ __forceinline__ Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return Dual(r, a.d / (2.0 * r));
}
__host__ __device__ __forceinline__ Dual sin(Dual a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
__host__ __device__ __forceinline__ Dual cos(Dual a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
__host__ __device__ __forceinline__ double val(double a) { return a; }
__host__ __device__ __forceinline__ double val(Dual a) { return a.v; }

template <typename S>
struct V3T {
  S x, y, z;
};
using V3 = V3T<double>;
template <typename S>
__host__ __device__ __forceinline__ V3T<S> v3(S a, S b, S c) { return V3T<S>{a, b, c}; }
template <typename S>
__host__ __device__ __forceinline__ V3T<S> lift(V3 a) { return V3T<S>{S(a.x), S(a.y), S(a.z)}; }
template <typename S>
__host__ __device__ __forceinline__ V3T<S> operator+(V3T<S> a, V3T<S> b) { return v3<S>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename S>
__host__ __device__ __forceinline__ V3T<S> operator-(V3T<S> a, V3T<S> b) { return v3<S>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename S>
__host__ __device__ __forceinline__ V3T<S> operator*(S s, V3T<S> a) { return v3<S>(s * a.x, s * a.y, s * a.z); }
template <typename S>
__host__ __device__ __forceinline__ S dot(V3T<S> a, V3T<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename S>
__host__ __device__ __forceinline__ V3T<S> cross(V3T<S> a, V3T<S> b) {
  return v3<S>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename S>
__host__ __device__ __forceinline__ S norm(V3T<S> a) { return sqrt(dot(a, a)); }

struct BeamGeom {
  V3 X1, X2;
  double E, G, A, Iy, Iz, J;
};

__device__ __forceinline__ BeamGeom load_geom(int e, const int* __restrict__ conn, const double* __restrict__ nodes,
                                              const double* __restrict__ props) {
  const int n1 = conn[2 * e], n2 = conn[2 * e + 1];
  BeamGeom g;
  g.X1 = v3<double>(nodes[3 * n1], nodes[3 * n1 + 1], nodes[3 * n1 + 2]);
  g.X2 = v3<double>(nodes[3 * n2], nodes[3 * n2 + 1], nodes[3 * n2 + 2]);
  const double* p = props + 6 * (size_t)e;
  g.E = p[0];
  g.G = p[1];
  g.A = p[2];
  g.Iy = p[3];
  g.Iz = p[4];
  g.J = p[5];
  return g;
}

// f[12] = B^T q of one element for the element displacement u[12]
// ([u1 theta1 u2 theta2]).  S = double: the reference's forces.  S = Dual:
// the same forces and their derivative along the direction carried by u's
// dual parts.  The rotation taking e1 to r1 is the reference's Rodrigues
// form, cos(phi) e2 + sin(phi) (a x e2) + (a.e2)(1 - cos(phi)) a with
// a = (e1 x r1)/|e1 x r1| and phi = atan2(|e1 x r1|, e1.r1), skipped when
// |e1 x r1| <= 1e-12.  With c = e1 x r1, |c| = sin(phi), e1.r1 = cos(phi)
// (unit vectors) it equals cos(phi) e2 + c x e2 + (c.e2)/(1 + cos(phi)) c,
// which is smooth through c = 0: the differentiated form uses it, so the
// tangent at an unrotated chord is the derivative of the rotation there
// (the limit the reference's branch approaches), not the zero derivative
// of its skipped branch.
template <typename S>
__host__ __device__ void corot_force(const BeamGeom& g, const S* u, S* f) {
  const V3 v0 = g.X2 - g.X1;
  const double L0 = norm(v0);
  const V3 e1 = (1.0 / fmax(L0, 1e-15)) * v0;
  const V3 ref = fabs(e1.z) < 0.9 ? v3<double>(0.0, 0.0, 1.0) : v3<double>(0.0, 1.0, 0.0);
  V3 e2d = cross(ref, e1);
  e2d = (1.0 / norm(e2d)) * e2d;
  V3 e3d = cross(e1, e2d);
  e3d = (1.0 / norm(e3d)) * e3d;
  const V3T<S> e2 = lift<S>(e2d), e3 = lift<S>(e3d), E1 = lift<S>(e1);
  const V3T<S> u1 = v3<S>(u[0], u[1], u[2]), th1 = v3<S>(u[3], u[4], u[5]);
  const V3T<S> u2 = v3<S>(u[6], u[7], u[8]), th2 = v3<S>(u[9], u[10], u[11]);
  const V3T<S> du = u2 - u1;
  const V3T<S> v = lift<S>(v0) + du;
  const S L = norm(v);
  V3T<S> r1 = (S(1.0) / (val(L) > 1e-15 ? L : S(1e-15))) * v;
  const V3T<S> cr = cross(E1, r1);
  const S cphi = dot(E1, r1);
  V3T<S> r2 = e2, r3 = e3;
  if constexpr (std::is_same<S, double>::value) {
    const double sphi = norm(cr);
    if (sphi > 1e-12) {
      const V3 ax = (1.0 / sphi) * cr;
      const double phi = atan2(sphi, cphi);
      const double cp = cos(phi), sp = sin(phi);
      r2 = cp * e2 + sp * cross(ax, e2) + (dot(ax, e2) * (1.0 - cp)) * ax;
      r3 = cp * e3 + sp * cross(ax, e3) + (dot(ax, e3) * (1.0 - cp)) * ax;
    }
  } else {
    const S k = S(1.0) / (S(1.0) + cphi);
    r2 = cphi * e2 + cross(cr, e2) + (dot(cr, e2) * k) * cr;
    r3 = cphi * e3 + cross(cr, e3) + (dot(cr, e3) * k) * cr;
  }
  const S tw = S(0.5) * dot(th1 + th2, r1);
  const S ct = cos(tw), st = sin(tw);
  const V3T<S> r2t = ct * r2 + st * cross(r1, r2);
  r1 = (S(1.0) / norm(r1)) * r1;
  r2 = r2t - dot(r2t, r1) * r1;
  r2 = (S(1.0) / norm(r2)) * r2;
  r3 = cross(r1, r2);  // the reference recomputes r3 from r1 x r2 after the twist
  r3 = (S(1.0) / norm(r3)) * r3;
  const V3T<S> V0 = lift<S>(v0);
  const S ul = (S(2.0) * dot(V0, du) + dot(du, du)) / (L + S(L0));
  const S iL = S(1.0) / L;
  const S thy1 = dot(r2, th1) + dot(r3, du) * iL;
  const S thz1 = dot(r3, th1) - dot(r2, du) * iL;
  const S thx = dot(r1, th2 - th1);
  const S thy2 = dot(r2, th2) + dot(r3, du) * iL;
  const S thz2 = dot(r3, th2) - dot(r2, du) * iL;
  const S N = S(g.E * g.A / L0) * ul;
  const S ky = S(g.E * g.Iy / L0), kz = S(g.E * g.Iz / L0);
  const S My1 = ky * (S(4.0) * thy1 + S(2.0) * thy2);
  const S My2 = ky * (S(2.0) * thy1 + S(4.0) * thy2);
  const S Mz1 = kz * (S(4.0) * thz1 + S(2.0) * thz2);
  const S Mz2 = kz * (S(2.0) * thz1 + S(4.0) * thz2);
  const S Tx = S(g.G * g.J / L0) * thx;
  const V3T<S> f1 = (-N) * r1 - ((My1 + My2) * iL) * r3 + ((Mz1 + Mz2) * iL) * r2;
  const V3T<S> m1 = (-Tx) * r1 + My1 * r2 + Mz1 * r3;
  const V3T<S> m2 = Tx * r1 + My2 * r2 + Mz2 * r3;
  f[0] = f1.x;
  f[1] = f1.y;
  f[2] = f1.z;
  f[3] = m1.x;
  f[4] = m1.y;
  f[5] = m1.z;
  f[6] = -f1.x;
  f[7] = -f1.y;
  f[8] = -f1.z;
  f[9] = m2.x;
  f[10] = m2.y;
  f[11] = m2.z;
}

__device__ __forceinline__ void load_u(int e, const int* __restrict__ conn, const double* __restrict__ U, double* u) {
  const int n1 = conn[2 * e], n2 = conn[2 * e + 1];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    u[k] = U[6 * (size_t)n1 + k];
    u[6 + k] = U[6 * (size_t)n2 + k];
  }
}

__global__ void beam_force_kernel(int M, const int* __restrict__ conn, const double* __restrict__ nodes,
                                  const double* __restrict__ props, const double* __restrict__ U,
                                  double* __restrict__ fe) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  const BeamGeom g = load_geom(e, conn, nodes, props);
  double u[12], f[12];
  load_u(e, conn, U, u);
  corot_force<double>(g, u, f);
#pragma unroll
  for (int k = 0; k < 12; ++k) fe[12 * (size_t)e + k] = f[k];
}

// Column j of the element tangent d f / d u_j, one thread per (element,
// column).  h_rel <= 0: the exact derivative (forces evaluated once in dual
// numbers seeded with e_j).  h_rel > 0: central differences
// Kt[e][i][j] = (f_i(u + h e_j) - f_i(u - h e_j)) / 2h, h = h_rel * L0 for
// translations and h_rel for rotations.
__global__ void beam_tangent_kernel(int M, const int* __restrict__ conn, const double* __restrict__ nodes,
                                    const double* __restrict__ props, const double* __restrict__ U, double h_rel,
                                    double* __restrict__ Kt) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 12LL * M) return;
  const int e = (int)(t / 12), j = (int)(t - 12LL * e);
  const BeamGeom g = load_geom(e, conn, nodes, props);
  double u[12];
  load_u(e, conn, U, u);
  double* out = Kt + 144 * (size_t)e;
  if (h_rel <= 0.0) {
    Dual ud[12], fd[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) ud[k] = Dual(u[k], k == j ? 1.0 : 0.0);
    corot_force<Dual>(g, ud, fd);
#pragma unroll
    for (int i = 0; i < 12; ++i) out[12 * i + j] = fd[i].d;
    return;
  }
  double fp[12], fm[12];
  const double L0 = norm(g.X2 - g.X1);
  const double h = ((j % 6) < 3 ? h_rel * L0 : h_rel);
  const double uj = u[j];
  u[j] = uj + h;
  corot_force<double>(g, u, fp);
  u[j] = uj - h;
  corot_force<double>(g, u, fm);
  const double inv = 0.5 / h;
#pragma unroll
  for (int i = 0; i < 12; ++i) out[12 * i + j] = (fp[i] - fm[i]) * inv;
}

// numpy.isclose(a, b) with its default tolerances (rtol 1e-5, atol 1e-8)
__device__ __forceinline__ bool isclose(double a, double b) { return fabs(a - b) <= 1e-8 + 1e-5 * fabs(b); }

// Entry (a, b) of the 12x12 local linear beam stiffness (assembler.py:37-84).
__device__ double ke_local(int a, int b, double L, const BeamGeom& g) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  const double ax = g.E * g.A / L, tor = g.G * g.J / L;
  const double bz1 = 12 * g.E * g.Iz / (L * L * L), bz2 = 6 * g.E * g.Iz / (L * L), bz3 = 4 * g.E * g.Iz / L,
               bz4 = 2 * g.E * g.Iz / L;
  const double by1 = 12 * g.E * g.Iy / (L * L * L), by2 = 6 * g.E * g.Iy / (L * L), by3 = 4 * g.E * g.Iy / L,
               by4 = 2 * g.E * g.Iy / L;
  switch (a * 12 + b) {
    case 0 * 12 + 0: case 6 * 12 + 6: return ax;
    case 0 * 12 + 6: return -ax;
    case 3 * 12 + 3: case 9 * 12 + 9: return tor;
    case 3 * 12 + 9: return -tor;
    case 1 * 12 + 1: case 7 * 12 + 7: return bz1;
    case 1 * 12 + 7: return -bz1;
    case 1 * 12 + 5: case 1 * 12 + 11: return bz2;
    case 5 * 12 + 5: case 11 * 12 + 11: return bz3;
    case 5 * 12 + 11: return bz4;
    case 5 * 12 + 7: case 7 * 12 + 11: return -bz2;
    case 2 * 12 + 2: case 8 * 12 + 8: return by1;
    case 2 * 12 + 8: return -by1;
    case 2 * 12 + 4: case 2 * 12 + 10: return -by2;
    case 4 * 12 + 4: case 10 * 12 + 10: return by3;
    case 4 * 12 + 10: return by4;
    case 4 * 12 + 8: case 8 * 12 + 10: return by2;
    default: return 0.0;
  }
}

// Row r of T^T K_local T per thread (assembler.py:87-148 frame rules).
__global__ void beam_ke_linear_kernel(int M, const int* __restrict__ conn, const double* __restrict__ nodes,
                                      const double* __restrict__ props, double* __restrict__ Ke) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= 12LL * M) return;
  const int e = (int)(t / 12), r = (int)(t - 12LL * e);
  const BeamGeom g = load_geom(e, conn, nodes, props);
  const V3 vec = g.X2 - g.X1;
  const double L = norm(vec);
  double* out = Ke + 144 * (size_t)e + 12 * r;
  if (L < 1e-12) {
#pragma unroll
    for (int c = 0; c < 12; ++c) out[c] = 0.0;
    return;
  }
  double R[3][3];
  V3 lx, ly, lz;
  if (isclose(fabs(vec.x), L)) {
    lx = v3(vec.x > 0 ? 1.0 : -1.0, 0.0, 0.0);
    ly = v3(0.0, 1.0, 0.0);
    lz = v3(0.0, 0.0, 1.0);
  } else if (isclose(fabs(vec.y), L)) {
    lx = v3(0.0, vec.y > 0 ? 1.0 : -1.0, 0.0);
    ly = v3(1.0, 0.0, 0.0);
    lz = v3(0.0, 0.0, 1.0);
  } else if (isclose(fabs(vec.z), L)) {
    lx = v3(0.0, 0.0, vec.z > 0 ? 1.0 : -1.0);
    ly = v3(1.0, 0.0, 0.0);
    lz = v3(0.0, 1.0, 0.0);
  } else {
    lx = (1.0 / L) * vec;
    ly = cross(v3(0.0, 0.0, 1.0), lx);
    if (norm(ly) < 1e-9) ly = cross(v3(0.0, 1.0, 0.0), lx);
    ly = (1.0 / norm(ly)) * ly;
    lz = cross(lx, ly);
    lz = (1.0 / norm(lz)) * lz;
  }
  R[0][0] = lx.x; R[0][1] = lx.y; R[0][2] = lx.z;
  R[1][0] = ly.x; R[1][1] = ly.y; R[1][2] = ly.z;
  R[2][0] = lz.x; R[2][1] = lz.y; R[2][2] = lz.z;
  // (T^T K T)[r][c] = sum_{i,j} R[i][r%3] K[3I+i][3J+j] R[j][c%3], I = r/3, J = c/3
  const int I = r / 3, ri = r % 3;
  for (int c = 0; c < 12; ++c) {
    const int Jb = c / 3, ci = c % 3;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s += R[i][ri] * ke_local(3 * I + i, 3 * Jb + j, L, g) * R[j][ci];
    out[c] = s;
  }
}

// out[n][0..5] (+)= sum of the element forces of the elements at node n
// (incidence codes e*2 + local, in element order).
__global__ void gather6_kernel(int n_nodes, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                               const double* __restrict__ fe, double* __restrict__ out, int accumulate) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; ++k) {
    const int code = n2e[k];
    const double* p = fe + 12 * (size_t)(code >> 1) + 6 * (code & 1);
#pragma unroll
    for (int c = 0; c < 6; ++c) s[c] += p[c];
  }
  double* o = out + 6 * (size_t)n;
#pragma unroll
  for (int c = 0; c < 6; ++c) o[c] = accumulate ? o[c] + s[c] : s[c];
}

int launch_beam_forces(int M, int n_nodes, const int* conn, const double* nodes, const double* props,
                       const double* U, const int* n2e_ptr, const int* n2e, double* fe, double* F, int accumulate,
                       hipStream_t s) {
  if (M > 0) {
    hipLaunchKernelGGL(beam_force_kernel, dim3(ceil_div(M, 128)), dim3(128), 0, s, M, conn, nodes, props, U, fe);
    const int e = static_cast<int>(hipGetLastError());
    if (e) return e;
  }
  if (n_nodes <= 0) return kOk;
  hipLaunchKernelGGL(gather6_kernel, dim3(ceil_div(n_nodes, 256)), dim3(256), 0, s, n_nodes, n2e_ptr, n2e, fe, F,
                     accumulate);
  return static_cast<int>(hipGetLastError());
}

int launch_beam_tangent(int M, const int* conn, const double* nodes, const double* props, const double* U,
                        double h_rel, double* Kt, hipStream_t s) {
  if (M <= 0) return kOk;
  hipLaunchKernelGGL(beam_tangent_kernel, dim3(ceil_div(12LL * M, 128)), dim3(128), 0, s, M, conn, nodes, props, U,
                     h_rel, Kt);
  return static_cast<int>(hipGetLastError());
}

int launch_beam_ke_linear(int M, const int* conn, const double* nodes, const double* props, double* Ke,
                          hipStream_t s) {
  if (M <= 0) return kOk;
  hipLaunchKernelGGL(beam_ke_linear_kernel, dim3(ceil_div(12LL * M, 128)), dim3(128), 0, s, M, conn, nodes, props,
                     Ke);
  return static_cast<int>(hipGetLastError());
}

}  // namespace wn
