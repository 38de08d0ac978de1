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
