// Fields computed on the mesh: element sizes, the size-field producers of
// the reference's warp loop (test_warp.cpp:70-83: element gradients of a
// vertex field, volume-weighted recovery to the vertices, twice, then a
// size from the Hessian's norm) and the identity size field.  One thread
// per element or vertex; the FP64 operations are the reference's, in its
// order, so results match bit for bit.
#include <cmath>
#include <string>

#include "ohx_geom.hpp"
#include "ohx_mesh.hpp"
#include "ohx_prims.hpp"

namespace ohx {

namespace {

// element_sizes (size.cpp:14-55): volume or area
__global__ __launch_bounds__(BLOCK) void element_sizes_k(u32 n, u32 dim, u32 const* __restrict__ ev,
                                                         double const* __restrict__ x, double* out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 const* v = ev + (size_t) i * (dim + 1);
  if (dim == 3)
    out[i] = tet_volume(ld3(x + 3 * (size_t) v[0]), ld3(x + 3 * (size_t) v[1]), ld3(x + 3 * (size_t) v[2]),
                        ld3(x + 3 * (size_t) v[3]));
  else
    out[i] = tri_area(ld3(x + 3 * (size_t) v[0]), ld3(x + 3 * (size_t) v[1]), ld3(x + 3 * (size_t) v[2]));
}

// element_jacobian + invert_jacobian (jacobian.cpp:7-66; invert_2x2 and
// invert_3x3, algebra.hpp:171-196): the inverse (3D) or pseudo-inverse
// (2D) of the matrix whose column i is the edge from vertex 0 to i + 1
__device__ inline void inverse_jacobian(u32 dim, V3 const* x, double out[3][3]) {
  double in[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (u32 i = 0; i < dim; ++i) {
    in[0][i] = x[i + 1].x - x[0].x;
    in[1][i] = x[i + 1].y - x[0].y;
    in[2][i] = x[i + 1].z - x[0].z;
  }
  if (dim == 3) {
    V3 r0{in[0][0], in[0][1], in[0][2]}, r1{in[1][0], in[1][1], in[1][2]}, r2{in[2][0], in[2][1], in[2][2]};
    double d = dot(r2, cross(r0, r1));
    V3 t0 = cross(r1, r2), t1 = cross(r2, r0), t2 = cross(r0, r1);
    double s = 1.0 / d;
    out[0][0] = t0.x * s;
    out[0][1] = t1.x * s;
    out[0][2] = t2.x * s;
    out[1][0] = t0.y * s;
    out[1][1] = t1.y * s;
    out[1][2] = t2.y * s;
    out[2][0] = t0.z * s;
    out[2][1] = t1.z * s;
    out[2][2] = t2.z * s;
    return;
  }
  double nrm[2][2];
  for (u32 i = 0; i < 2; ++i)
    for (u32 j = 0; j < 2; ++j) {
      double a = 0;
      for (u32 k = 0; k < 3; ++k) a += in[i][k] * in[j][k];
      nrm[i][j] = a;
    }
  double det = nrm[0][0] * nrm[1][1] - nrm[0][1] * nrm[1][0];
  double inv[2][2] = {{nrm[1][1] / det, -nrm[0][1] / det}, {-nrm[1][0] / det, nrm[0][0] / det}};
  for (u32 i = 0; i < 3; ++i)
    for (u32 j = 0; j < 3; ++j) out[i][j] = 0;
  for (u32 i = 0; i < 3; ++i)
    for (u32 j = 0; j < 2; ++j)
      for (u32 k = 0; k < 2; ++k) out[i][j] += in[k][i] * inv[k][j];
}

// execute (element_gradients.cpp:15-46)
__global__ __launch_bounds__(BLOCK) void element_gradients_k(u32 n, u32 dim, u32 const* __restrict__ ev,
                                                             double const* __restrict__ x, u32 nc,
                                                             double const* __restrict__ f, double* out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 vpe = dim + 1;
  V3 ex[4];
  u32 vs[4];
  for (u32 j = 0; j < vpe; ++j) {
    vs[j] = ev[(size_t) i * vpe + j];
    ex[j] = ld3(x + 3 * (size_t) vs[j]);
  }
  double ji[3][3];
  inverse_jacobian(dim, ex, ji);
  double* g = out + (size_t) i * nc * 3;
  for (u32 k = 0; k < nc; ++k) {
    double f0 = f[(size_t) vs[0] * nc + k];
    double acc[3] = {0, 0, 0};
    for (u32 j = 0; j < dim; ++j) {
      double df = f[(size_t) vs[j + 1] * nc + k] - f0;
      for (u32 l = 0; l < 3; ++l) acc[l] += ji[j][l] * df;
    }
    for (u32 l = 0; l < 3; ++l) g[k * 3 + l] = acc[l];
  }
}

// recover_vert_by_volume (recover_by_volume.cpp:10-33)
__global__ __launch_bounds__(BLOCK) void recover_by_volume_k(u32 nv, u32 const* __restrict__ uo,
                                                             u32 const* __restrict__ ua,
                                                             double const* __restrict__ sizes, u32 nc,
                                                             double const* __restrict__ f, double* out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nv) return;
  double* o = out + (size_t) i * nc;
  for (u32 k = 0; k < nc; ++k) {
    double ssum = 0, acc = 0;
    for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
      u32 el = ua[j];
      double es = sizes[el];
      ssum += es;
      acc += es * f[(size_t) el * nc + k];
    }
    o[k] = acc / ssum;
  }
}

// vert_size_from_hessian + clamp (size_from_hessian.cpp:11-41)
__global__ __launch_bounds__(BLOCK) void size_from_hessian_k(u32 nv, u32 nh, double const* __restrict__ h,
                                                             double const* __restrict__ w, double min_h,
                                                             double max_h, double* out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nv) return;
  double const* hv = h + (size_t) i * nh;
  double total = 0;
  for (u32 j = 0; j < nh / 9; ++j) {
    double d = 0;
    for (u32 k = 0; k < 9; ++k) d += hv[k] * hv[k];
    double nrm = sqrt(d);
    total += (w ? w[j] : 1.0) * nrm;
    hv += 9;
  }
  double s = max_h - total;
  if (s < min_h) s = min_h;
  out[i] = s;
}

}  // namespace

Buf<double> element_sizes(Mesh& m) {
  u32 n = count(m, m.dim);
  Buf<double> out(n);
  if (n) {
    hipLaunchKernelGGL(element_sizes_k, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, m.dim, ask_down(m, m.dim, 0),
                       coords(m), out.get());
    check_launch("element_sizes");
  }
  return out;
}

// mesh_element_gradients (element_gradients.cpp:66-89)
void element_gradients(Mesh& m, char const* name) {
  Tag* t = find_tag(m, 0, name);
  OHX_REQUIRE(t && t->type == TagType::F64, "element_gradients: no such vertex field");
  OHX_REQUIRE(m.dim == 2 || m.dim == 3, "element_gradients: 2D or 3D mesh");
  u32 n = count(m, m.dim), nc = t->ncomps;
  Buf<double> out((size_t) n * nc * 3);
  if (n) {
    hipLaunchKernelGGL(element_gradients_k, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, m.dim,
                       ask_down(m, m.dim, 0), coords(m), nc, t->f.get(), out.get());
    check_launch("element_gradients");
  }
  add_tag_f64(m, m.dim, ("grad_" + std::string(name)).c_str(), nc * 3, std::move(out));
}

// mesh_recover_by_volume (recover_by_volume.cpp:52-66)
void recover_by_volume(Mesh& m, char const* name) {
  Tag* t = find_tag(m, m.dim, name);
  OHX_REQUIRE(t && t->type == TagType::F64, "recover_by_volume: no such element field");
  u32 nv = count(m, 0), nc = t->ncomps;
  Buf<double> sizes = element_sizes(m);
  Up const& up = ask_up(m, 0, m.dim);
  Buf<double> out((size_t) nv * nc);
  if (nv) {
    hipLaunchKernelGGL(recover_by_volume_k, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, up.offsets.get(),
                       up.adj.get(), sizes.get(), nc, t->f.get(), out.get());
    check_launch("recover_by_volume");
  }
  add_tag_f64(m, 0, name, nc, std::move(out));
}

// mesh_size_from_hessian (size_from_hessian.cpp:44-72): vertex field
// "adapt_size"; weights (one per solution component) may be null
void size_from_hessian(Mesh& m, char const* hess_name, double const* weights, double min_h, double max_h) {
  Tag* t = find_tag(m, 0, hess_name);
  OHX_REQUIRE(t && t->type == TagType::F64 && t->ncomps % 9 == 0,
              "size_from_hessian: needs a vertex field of 9 components per solution component");
  OHX_REQUIRE(max_h > min_h && min_h > 0, "size_from_hessian: needs 0 < min_h < max_h");
  u32 ns = t->ncomps / 9;
  if (weights)
    for (u32 j = 0; j < ns; ++j) OHX_REQUIRE(weights[j] >= 0, "size_from_hessian: negative weight");
  u32 nv = count(m, 0);
  Buf<double> w;
  if (weights) w = to_device(weights, ns);
  Buf<double> out(nv);
  if (nv) {
    hipLaunchKernelGGL(size_from_hessian_k, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, t->ncomps, t->f.get(),
                       weights ? w.get() : nullptr, min_h, max_h, out.get());
    check_launch("size_from_hessian");
  }
  add_tag_f64(m, 0, "adapt_size", 1, std::move(out));
}

// mesh_identity_size_field (size.cpp:90-135): longest adjacent edge
void identity_size(Mesh& m, char const* name) {
  Graph const& g = ask_star(m, 0, 1);
  u32 nv = count(m, 0);
  double const* x = coords(m);
  Buf<double> out(nv);
  double* o = out.get();
  u32 const* go = g.offsets.get();
  u32 const* ga = g.adj.get();
  parallel_for(nv, OHX_LAMBDA(u32 i) {
    V3 a = ld3(x + 3 * (size_t) i);
    double mx = 0;
    for (u32 j = go[i]; j < go[i + 1]; ++j) {
      double l = sqrt(sqdist(ld3(x + 3 * (size_t) ga[j]), a));
      if (l > mx) mx = l;
    }
    o[i] = mx;
  });
  add_tag_f64(m, 0, name, 1, std::move(out));
}

}  // namespace ohx
