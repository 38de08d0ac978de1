// FP64 geometry and quality, in the reference's exact operation order.
//
// The library is compiled with -ffp-contract=off: no multiply-add is fused,
// so every operation rounds as in the reference's x86-64 SSE2 build, and
// sqrt / division are IEEE correctly rounded on gfx950.  Ties in the adapt
// decisions (indset goodness, collapse targets, swap codes) are common on
// structured meshes, so these must be bit-identical, not merely close.
#pragma once

#include "ohx_common.hpp"

namespace ohx {

struct V3 {
  double x, y, z;
};

__host__ __device__ inline V3 ld3(double const* p) { return V3{p[0], p[1], p[2]}; }
__host__ __device__ inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// algebra.hpp:66-84
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double dot(V3 a, V3 b) {
  double d = 0;
  d += a.x * b.x;
  d += a.y * b.y;
  d += a.z * b.z;
  return d;
}
// vector_squared_distance (algebra.hpp:92-104)
__host__ __device__ inline double sqdist(V3 a, V3 b) {
  double s = 0, d;
  d = a.x - b.x;
  s += d * d;
  d = a.y - b.y;
  s += d * d;
  d = a.z - b.z;
  s += d * d;
  return s;
}

// size.hpp:14-46
__host__ __device__ inline double tri_area(V3 a, V3 b, V3 c) {
  V3 x = cross(sub(b, a), sub(c, a));
  return sqrt(dot(x, x)) / 2.0;
}
__host__ __device__ inline double tri_z_area(V3 a, V3 b, V3 c) {
  V3 x = cross(sub(b, a), sub(c, a));
  return x.z / 2.0;
}
__host__ __device__ inline double tet_volume(V3 a, V3 b, V3 c, V3 d) {
  V3 x = cross(sub(b, a), sub(c, a));
  return dot(x, sub(d, a)) / 6.0;
}

// quality.hpp:15-18, computed in the same order
__host__ __device__ inline double perfect_tri() { return sqrt(3.0) / 4.0; }
__host__ __device__ inline double perfect_tet() {
  double s = sqrt(perfect_tri());
  return (sqrt(2.0) / 12.0) / (s * s * s);
}

// triangle_quality / triangle_xy_quality (quality.hpp:34-66)
__host__ __device__ inline double tri_quality(V3 a, V3 b, V3 c, bool signed_z) {
  double sum_lsq = 0;
  sum_lsq += sqdist(a, b);
  sum_lsq += sqdist(b, c);
  sum_lsq += sqdist(c, a);
  double lrms = sqrt(sum_lsq / 3);
  double ar = signed_z ? tri_z_area(a, b, c) : tri_area(a, b, c);
  double q = ar / (lrms * lrms);
  return q / perfect_tri();
}

// tet_quality (quality.hpp:68-89): face order {0,1,2},{1,0,3},{2,1,3},{0,2,3}
__host__ __device__ inline double tet_quality(V3 x0, V3 x1, V3 x2, V3 x3) {
  double sum_asq = 0, a;
  a = tri_area(x0, x1, x2);
  sum_asq += a * a;
  a = tri_area(x1, x0, x3);
  sum_asq += a * a;
  a = tri_area(x2, x1, x3);
  sum_asq += a * a;
  a = tri_area(x0, x2, x3);
  sum_asq += a * a;
  double arms = sqrt(sum_asq / 4.);
  double v = tet_volume(x0, x1, x2, x3);
  double r = sqrt(arms);
  double q = v / (r * r * r);
  return q / perfect_tet();
}

// element_quality (signed 2D area) vs entity_quality (unsigned), quality.hpp:91-109
__host__ __device__ inline double element_quality(u32 dim, V3 const* x) {
  return dim == 3 ? tet_quality(x[0], x[1], x[2], x[3]) : tri_quality(x[0], x[1], x[2], true);
}
__host__ __device__ inline double entity_quality(u32 dim, V3 const* x) {
  return dim == 3 ? tet_quality(x[0], x[1], x[2], x[3]) : tri_quality(x[0], x[1], x[2], false);
}

}  // namespace ohx

namespace ohx {

// ------------------------------------------------------------------------
// Anisotropic metric extension (no counterpart in the reference, which has
// only the scalar "adapt_size").  A vertex field "metric" holds a symmetric
// positive definite 3x3 tensor as (m00, m01, m02, m11, m12, m22).  An
// element's metric is the average of its vertices' (summed in vertex order,
// then divided, like average_element_field); its quality is the reference's
// quality formula evaluated on the edge vectors mapped by the Cholesky
// factor U (U^T U = M), so it is the v1 measure in metric space; an edge's
// metric length is sqrt(e^T M e) with M the average of its two vertices'.
// FP64 throughout, fixed operation order (the oracle repeats it exactly).
struct Sym {
  double m[6];
};

__host__ __device__ inline Sym ld6(double const* p) {
  Sym s;
  for (int k = 0; k < 6; ++k) s.m[k] = p[k];
  return s;
}

template <int N>
__host__ __device__ inline Sym sym_average(Sym const* v) {
  Sym s;
  for (int k = 0; k < 6; ++k) {
    double a = 0;
    for (int i = 0; i < N; ++i) a += v[i].m[k];
    a /= N;
    s.m[k] = a;
  }
  return s;
}

// upper Cholesky factor: u00 u01 u02 / u11 u12 / u22
struct Chol {
  double u00, u01, u02, u11, u12, u22;
};

__host__ __device__ inline Chol cholesky(Sym const& a) {
  Chol c;
  c.u00 = sqrt(a.m[0]);
  c.u01 = a.m[1] / c.u00;
  c.u02 = a.m[2] / c.u00;
  c.u11 = sqrt(a.m[3] - c.u01 * c.u01);
  c.u12 = (a.m[4] - c.u01 * c.u02) / c.u11;
  c.u22 = sqrt((a.m[5] - c.u02 * c.u02) - c.u12 * c.u12);
  return c;
}

__host__ __device__ inline V3 chol_map(Chol const& c, V3 e) {
  return V3{(c.u00 * e.x + c.u01 * e.y) + c.u02 * e.z, c.u11 * e.y + c.u12 * e.z, c.u22 * e.z};
}

// e^T M e for the average M of two vertex metrics
__host__ __device__ inline double metric_len2(Sym const& a, Sym const& b, V3 e) {
  double m[6];
  for (int k = 0; k < 6; ++k) {
    double s = 0;
    s += a.m[k];
    s += b.m[k];
    s /= 2;
    m[k] = s;
  }
  double s = 0;
  s += e.x * ((m[0] * e.x + m[1] * e.y) + m[2] * e.z);
  s += e.y * ((m[1] * e.x + m[3] * e.y) + m[4] * e.z);
  s += e.z * ((m[2] * e.x + m[4] * e.y) + m[5] * e.z);
  return s;
}

// quality of a simplex with per-vertex metrics (dim 2 or 3); signed_z picks
// the signed z-area triangle form (element quality) over the unsigned one
template <int DIM>
__host__ __device__ inline double metric_quality(V3 const* x, Sym const* mv, bool signed_z) {
  Sym me = sym_average<DIM + 1>(mv);
  Chol c = cholesky(me);
  V3 o{0.0, 0.0, 0.0};
  V3 p1 = chol_map(c, sub(x[1], x[0]));
  V3 p2 = chol_map(c, sub(x[2], x[0]));
  if (DIM == 3) {
    V3 p3 = chol_map(c, sub(x[3], x[0]));
    return tet_quality(o, p1, p2, p3);
  }
  return tri_quality(o, p1, p2, signed_z);
}

}  // namespace ohx

namespace ohx {

// edge size for the adapt decisions: v1's |x1 - x0| / ((h0 + h1) / 2)
// (size.cpp:57-69), or the metric length sqrt(e^T M e) when mt != null
__host__ __device__ inline double edge_size(u32 a, u32 b, double const* x, double const* h, double const* mt) {
  V3 xa = ld3(x + 3 * (size_t) a), xb = ld3(x + 3 * (size_t) b);
  if (mt) return sqrt(metric_len2(ld6(mt + 6 * (size_t) a), ld6(mt + 6 * (size_t) b), sub(xb, xa)));
  double len = sqrt(sqdist(xb, xa));
  double want = (h[a] + h[b]) / 2;
  return len / want;
}

}  // namespace ohx

namespace ohx {

// ------------------------------------------------------------------------
// Metric from a Hessian (extension of size_from_hessian.cpp:12-40, which in
// the reference builds only a scalar size): symmetric eigendecomposition
// H = V diag(l) V^T by cyclic Jacobi rotations in FP64, eigenvalues mapped
// to l' = clamp(|l| / eps, 1/hmax^2, 1/hmin^2), M = V diag(l') V^T.
// Fixed operation order; the oracle repeats it exactly.
__host__ __device__ inline void jacobi_eigen3(double const* h6, double* lam, double (*V)[3]) {
  double A[3][3] = {{h6[0], h6[1], h6[2]}, {h6[1], h6[3], h6[4]}, {h6[2], h6[4], h6[5]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  int const P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0;
    off += A[0][1] * A[0][1];
    off += A[0][2] * A[0][2];
    off += A[1][2] * A[1][2];
    double dia = 0;
    dia += A[0][0] * A[0][0];
    dia += A[1][1] * A[1][1];
    dia += A[2][2] * A[2][2];
    if (off <= 1e-40 * dia || off == 0.0) break;
    for (int r = 0; r < 3; ++r) {
      int p = P[r], q = Q[r];
      double apq = A[p][q];
      if (apq == 0.0) continue;
      double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
      double at = theta < 0 ? -theta : theta;
      double t = 1.0 / (at + sqrt(theta * theta + 1.0));
      if (theta < 0) t = -t;
      double c = 1.0 / sqrt(t * t + 1.0);
      double s = t * c;
      double app = A[p][p], aqq = A[q][q];
      A[p][p] = app - t * apq;
      A[q][q] = aqq + t * apq;
      A[p][q] = 0.0;
      A[q][p] = 0.0;
      int k = 3 - p - q;
      double akp = A[k][p], akq = A[k][q];
      A[k][p] = c * akp - s * akq;
      A[p][k] = A[k][p];
      A[k][q] = s * akp + c * akq;
      A[q][k] = A[k][q];
      for (int i = 0; i < 3; ++i) {
        double vip = V[i][p], viq = V[i][q];
        V[i][p] = c * vip - s * viq;
        V[i][q] = s * vip + c * viq;
      }
    }
  }
  lam[0] = A[0][0];
  lam[1] = A[1][1];
  lam[2] = A[2][2];
}

__host__ __device__ inline void metric_from_hessian(double const* h6, double eps, double hmin, double hmax,
                                                    double* m6) {
  double lam[3], V[3][3];
  jacobi_eigen3(h6, lam, V);
  double lo = 1.0 / (hmax * hmax), hi = 1.0 / (hmin * hmin);
  double l[3];
  for (int k = 0; k < 3; ++k) {
    double a = lam[k] < 0 ? -lam[k] : lam[k];
    a = a / eps;
    if (a < lo) a = lo;
    if (a > hi) a = hi;
    l[k] = a;
  }
  int const I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2};
  for (int c = 0; c < 6; ++c) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += (V[I[c]][k] * l[k]) * V[J[c]][k];
    m6[c] = s;
  }
}

}  // namespace ohx

namespace ohx {

// ------------------------------------------------------------------------
// FP32 screening (the FP32-mixed precision mode).  Edge vectors are formed
// in FP64 (so no cancellation error reaches FP32) and everything after in
// FP32 with the hardware's fast square root; the result differs from the
// FP64 value by far less than SCREEN_MARGIN for any element the adapt loop
// meets, and every decision whose FP32 input lies within the margin of its
// threshold is re-evaluated in FP64.
constexpr float SCREEN_MARGIN = 1e-4f;

struct F3 {
  float x, y, z;
};
__device__ inline F3 f3(V3 a, V3 b) { return F3{(float) (a.x - b.x), (float) (a.y - b.y), (float) (a.z - b.z)}; }
__device__ inline F3 fsub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 fcross(F3 a, F3 b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// the v1 measures from the edge vectors e1 = x1 - x0, e2 = x2 - x0 (, e3):
// triangle  q = A / (L_rms^2) / (sqrt(3)/4),  tet  q = V / A_rms^{3/2} / c,
// with the hardware's single-precision square root and reciprocals (about
// one ulp each)
__device__ inline float tri_q32(F3 e1, F3 e2, bool signed_z) {
  F3 e12 = fsub(e2, e1);
  float s = fdot(e1, e1) + fdot(e12, e12) + fdot(e2, e2);
  F3 c = fcross(e1, e2);
  float a2 = signed_z ? c.z : __builtin_amdgcn_sqrtf(fdot(c, c));  // twice the area
  // A / (s / 3) / (sqrt(3) / 4) = (a2 / 2) * 12 / (sqrt(3) s)
  return a2 * 3.46410162f * __builtin_amdgcn_rcpf(s);
}
__device__ inline float tet_q32(F3 e1, F3 e2, F3 e3) {
  F3 c0 = fcross(e1, e2);                                   // face 0 1 2
  F3 c1 = fcross(F3{-e1.x, -e1.y, -e1.z}, fsub(e3, e1));    // face 1 0 3
  F3 c2 = fcross(fsub(e1, e2), fsub(e3, e2));               // face 2 1 3
  F3 c3 = fcross(e2, e3);                                   // face 0 2 3
  float s = fdot(c0, c0) + fdot(c1, c1) + fdot(c2, c2) + fdot(c3, c3);  // 4 * sum of A^2
  float vol6 = fdot(c0, e3);
  // A_rms^2 = s / 16, so A_rms^{3/2} = s^{3/4} / 8 and
  // q = (vol6 / 6) * 8 s^{-3/4} / perfect_tet = vol6 s^{-3/4} * 3.2237097
  float t = __builtin_amdgcn_rsqf(s);
  return vol6 * (t * __builtin_amdgcn_sqrtf(t)) * 3.22370979f;
}

// quality of a simplex from its FP64 vertices (and per-vertex metrics).
// The edge vectors, and with a metric the element metric's Cholesky factor
// and the mapped edge vectors, are formed in FP64 exactly as the FP64 path
// forms them: the factor of a strongly anisotropic metric is ill-conditioned
// in FP32 (an error of cond(M) * 2^-24), the mapped vectors are not.  Only the
// quality formula itself runs in FP32.
template <u32 DIM>
__device__ inline float quality32(V3 const* p, Sym const* pm, bool signed_z) {
  V3 d1 = sub(p[1], p[0]), d2 = sub(p[2], p[0]);
  V3 d3 = DIM == 3 ? sub(p[3], p[0]) : V3{0.0, 0.0, 0.0};
  if (pm) {
    Chol c = cholesky(sym_average<DIM + 1>(pm));
    d1 = chol_map(c, d1);
    d2 = chol_map(c, d2);
    d3 = chol_map(c, d3);
  }
  F3 e1{(float) d1.x, (float) d1.y, (float) d1.z}, e2{(float) d2.x, (float) d2.y, (float) d2.z};
  F3 e3{(float) d3.x, (float) d3.y, (float) d3.z};
  return DIM == 3 ? tet_q32(e1, e2, e3) : tri_q32(e1, e2, signed_z);
}

// screening statistics in device words: decisions settled in FP32, FP64
// re-evaluations, and the largest |q32 - q64| among re-evaluated values
// (float bits, so unsigned order is float order).  The counters are spread
// over STAT_SLOTS words 128 bytes apart, by block: atomics on one address
// retire one at a time in L2, so one shared counter would serialise every
// wave of the launch.
constexpr u32 STAT_SLOTS = 64;
constexpr u32 STAT_STRIDE = 16;  // u64 words between slots (128 bytes)
__device__ inline void wave_count(unsigned long long* w, bool pred) {
  unsigned long long m = __ballot(pred);
  unsigned long long act = __ballot(1);
  if (m && (u32) __lane_id() == (u32) (__ffsll((long long) act) - 1))
    atomicAdd(w + (size_t) (blockIdx.x % STAT_SLOTS) * STAT_STRIDE, (unsigned long long) __popcll(m));
}
__device__ inline void note_max_err(u32* w, float err) {
  // err >= 0: the atomic is skipped once a larger value is recorded
  u32 b = __float_as_uint(err);
  if (b > __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(w, b);
}

// element quality, isotropic (v1) or in the vertex metric when mt != null
template <u32 DIM>
__device__ inline double elem_q(u32 i, u32 const* __restrict__ ev, double const* __restrict__ x,
                                double const* __restrict__ mt) {
  V3 p[DIM + 1];
  u32 v[DIM + 1];
#pragma unroll
  for (u32 j = 0; j <= DIM; ++j) {
    v[j] = ev[(size_t) i * (DIM + 1) + j];
    p[j] = ld3(x + 3 * (size_t) v[j]);
  }
  if (!mt) return element_quality(DIM, p);
  Sym m[DIM + 1];
#pragma unroll
  for (u32 j = 0; j <= DIM; ++j) m[j] = ld6(mt + 6 * (size_t) v[j]);
  return metric_quality<DIM>(p, m, true);
}

}  // namespace ohx
