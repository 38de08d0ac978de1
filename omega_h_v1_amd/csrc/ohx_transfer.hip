// Solution transfer: element fields tagged OSH_TRANSFER_CONSERVE or
// OSH_TRANSFER_POINTWISE follow every refine, coarsen and swap
// (refine_conserve.cpp:10-71, refine_fit.cpp:10-70, coarsen_conserve.cpp:
// 15-139, coarsen_fit.cpp:17-187, swap_conserve.cpp:15-131, swap_fit.cpp:
// 16-127, qr.cpp:40-200); every other element tag is dropped by the
// operation.  Each output tag is the kept elements' values followed by the
// generated elements', in the operation's element order.  The arithmetic
// is the reference's, operation for operation, so the FP64 results match
// bit for bit:
//
// * conserve: a cavity's total is shared by its new elements in proportion
//   to their size (refine: equally among a split element's children);
// * pointwise: refine copies the parent value, swap takes the cavity mean,
//   coarsen fits a linear function of the element centroid by Householder
//   QR over the first 16 old elements of the cavity (16 lanes per cavity:
//   lane i holds row i of A, R and the explicit Q) and evaluates it at the
//   new centroids; fewer than 4 points or rank < 4 fall back to the mean.
#include <cmath>

#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_mesh.hpp"
#include "ohx_prims.hpp"

namespace ohx {

namespace {

constexpr int TR_POINTWISE = 1;
constexpr int TR_CONSERVE = 2;
constexpr u32 MAX_NCOMPS = 32;  // coarsen_conserve.cpp:11
constexpr u32 FIT_PTS = 16;     // MAX_PTS, qr.hpp:6; also the lanes per cavity

// up to 8 FP64 element fields handled by one launch
struct FieldList {
  u32 n = 0;
  u32 ncomps[8];
  double const* src[8];
  double* dst[8];
};

std::vector<Tag*> fields_of(Mesh& m, int kind) {
  std::vector<Tag*> out;
  for (auto& tp : m.tags[m.dim])
    if (tp->type == TagType::F64 && tp->transfer == kind) out.push_back(tp.get());
  return out;
}

// the output tags (in the input's order), as lists of at most 8
std::vector<FieldList> make_outputs(Mesh& m, Mesh& mo, std::vector<Tag*> const& ts, int kind, u32 nout) {
  std::vector<FieldList> out;
  for (Tag* t : ts) {
    if (out.empty() || out.back().n == 8) out.emplace_back();
    FieldList& L = out.back();
    Tag& nt = add_tag_f64(mo, m.dim, t->name.c_str(), t->ncomps, Buf<double>((size_t) nout * t->ncomps), kind);
    L.ncomps[L.n] = t->ncomps;
    L.src[L.n] = t->f.get();
    L.dst[L.n] = nt.f.get();
    ++L.n;
  }
  return out;
}

// ---------------------------------------------------------------- refine
// refine_conserve_elem / refine_fit_elem: survivors keep their values, the
// ppd children of a split element get value / ppd (conserve) or the value
__global__ __launch_bounds__(BLOCK) void refine_elem_fields(u32 n, u32 const* __restrict__ soff,
                                                            u32 const* __restrict__ off, u32 ppd, u32 go,
                                                            FieldList L, int conserve) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 k0 = off[i], k1 = off[i + 1];
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    double const* in = L.src[t];
    double* out = L.dst[t];
    if (k0 == k1) {
      u32 o = soff[i];
      for (u32 c = 0; c < w; ++c) out[(size_t) o * w + c] = in[(size_t) i * w + c];
      continue;
    }
    double denom = (double) (k1 * ppd - k0 * ppd);
    for (u32 j = 0; j < ppd; ++j) {
      size_t o = (size_t) go + (size_t) k0 * ppd + j;
      for (u32 c = 0; c < w; ++c) {
        double v = in[(size_t) i * w + c];
        out[o * w + c] = conserve ? v / denom : v;
      }
    }
  }
}

// ------------------------------------------------------- shared helpers
// the kept elements' values (expand_array over the survivor offsets)
__global__ __launch_bounds__(BLOCK) void copy_same(u32 n, u32 const* __restrict__ soff, FieldList L) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 o = soff[i];
  if (o == soff[i + 1]) return;
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    for (u32 c = 0; c < w; ++c) L.dst[t][(size_t) o * w + c] = L.src[t][(size_t) i * w + c];
  }
}

// average_element_field (element_field.hpp:8-25): ((0 + x0) + x1 ...) / vpe
__global__ __launch_bounds__(BLOCK) void centroids_k(u32 n, u32 vpe, u32 const* __restrict__ ev,
                                                     double const* __restrict__ x, double* out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  for (u32 c = 0; c < 3; ++c) {
    double s = 0;
    for (u32 k = 0; k < vpe; ++k) s += x[3 * (size_t) ev[(size_t) i * vpe + k] + c];
    s /= vpe;
    out[3 * (size_t) i + c] = s;
  }
}

// mesh_interp_to_elems(m, "coordinates") (element_field.cpp:36-45)
Tag& add_centroids(Mesh& m) {
  u32 n = count(m, m.dim);
  Buf<double> out((size_t) n * 3);
  if (n) {
    hipLaunchKernelGGL(centroids_k, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, m.dim + 1, ask_down(m, m.dim, 0),
                       coords(m), out.get());
    check_launch("element_centroids");
  }
  return add_tag_f64(m, m.dim, "coordinates", 3, std::move(out));
}

// ---------------------------------------------------------------- coarsen
// coarsen_conserve_cavity: one thread per collapsing vertex
__global__ __launch_bounds__(BLOCK) void coarsen_conserve_cavity(u32 ncol, u32 const* __restrict__ cols,
                                                                 u32 const* __restrict__ uo,
                                                                 u32 const* __restrict__ ua,
                                                                 u32 const* __restrict__ eg, u32 nsame,
                                                                 double const* __restrict__ sizes, FieldList L) {
  u32 c = blockIdx.x * BLOCK + threadIdx.x;
  if (c >= ncol) return;
  u32 v = cols[c];
  u32 f = uo[v], e = uo[v + 1];
  double cav = 0;
  for (u32 j = f; j < e; ++j) {
    u32 el = ua[j];
    if (eg[el] != eg[el + 1]) cav += sizes[eg[el] + nsame];
  }
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    double const* in = L.src[t];
    double* out = L.dst[t];
    for (u32 k = 0; k < w; ++k) {
      double s = 0;
      for (u32 j = f; j < e; ++j) s = s + in[(size_t) ua[j] * w + k];
      for (u32 j = f; j < e; ++j) {
        u32 el = ua[j];
        if (eg[el] == eg[el + 1]) continue;
        u32 ne = eg[el] + nsame;
        out[(size_t) ne * w + k] = s * (sizes[ne] / cav);
      }
    }
  }
}

__device__ inline double gshfl(double x, u32 src) { return __shfl(x, (int) src, (int) FIT_PTS); }

// coarsen_fit_cavity with qr_decomp2 / qr_solve2: 16 lanes per collapsing
// vertex, lane l owning row l of the fit (l < npts).  Every sum runs in
// the reference's order: values are gathered by shuffles and added from
// row k (or 0) upwards in each lane, so all lanes of a group agree.
__global__ __launch_bounds__(BLOCK) void coarsen_fit_cavity(u32 ncol, u32 const* __restrict__ cols,
                                                            u32 const* __restrict__ uo, u32 const* __restrict__ ua,
                                                            u32 const* __restrict__ eg, u32 nsame,
                                                            double const* __restrict__ oc,
                                                            double const* __restrict__ ncen, FieldList L) {
  u32 gid = (blockIdx.x * BLOCK + threadIdx.x) / FIT_PTS;
  u32 l = threadIdx.x % FIT_PTS;
  if (gid >= ncol) return;  // the whole group
  u32 v = cols[gid];
  u32 f = uo[v], e = uo[v + 1];
  u32 npts = e - f < FIT_PTS ? e - f : FIT_PTS;
  bool fit = npts >= 4;
  bool row = l < npts;
  double r[4] = {0, 0, 0, 0};
  double q[FIT_PTS];
#pragma unroll
  for (u32 j = 0; j < FIT_PTS; ++j) q[j] = (double) (l == j);
  double rr[4][4];
  if (fit) {
    if (row) {
      u32 el = ua[f + l];
      r[0] = 1.0;
      r[1] = oc[3 * (size_t) el];
      r[2] = oc[3 * (size_t) el + 1];
      r[3] = oc[3 * (size_t) el + 2];
    }
    u32 rank = 0;
#pragma unroll
    for (u32 k = 0; k < 4; ++k) {
      // get_reflector2
      double colk = r[k];
      double cn = 0;
#pragma unroll
      for (u32 i = 0; i < FIT_PTS; ++i) {
        double x = gshfl(colk, i);
        if (i >= k && i < npts) cn += x * x;
      }
      cn = sqrt(cn);
      if (cn < 1e-6) continue;
      double vl = (row && l >= k) ? colk : 0.0;
      double akk = gshfl(colk, k);
      if (l == k) vl += ((akk < 0) ? -1.0 : 1.0) * cn;
      double rn = 0;
#pragma unroll
      for (u32 i = 0; i < FIT_PTS; ++i) {
        double x = gshfl(vl, i);
        if (i >= k && i < npts) rn += x * x;
      }
      rn = sqrt(rn);
      if (row && l >= k) vl /= rn;
      // reflect_columns2
#pragma unroll
      for (u32 j = 0; j < 4; ++j) {
        double p = r[j] * vl;
        double dot = 0;
#pragma unroll
        for (u32 i = 0; i < FIT_PTS; ++i) {
          double x = gshfl(p, i);
          if (i >= k && i < npts) dot += x;
        }
        if (row && l >= k) r[j] -= 2 * dot * vl;
      }
      // reflect_rows2: every row l < npts against the whole reflector
      double vv[FIT_PTS];
#pragma unroll
      for (u32 i = 0; i < FIT_PTS; ++i) vv[i] = gshfl(vl, i);
      if (row) {
        double dot = 0;
#pragma unroll
        for (u32 j = 0; j < FIT_PTS; ++j)
          if (j >= k && j < npts) dot += q[j] * vv[j];
#pragma unroll
        for (u32 j = 0; j < FIT_PTS; ++j)
          if (j >= k && j < npts) q[j] -= 2 * dot * vv[j];
      }
      ++rank;
    }
    if (rank < 4) fit = false;
#pragma unroll
    for (u32 i = 0; i < 4; ++i)
#pragma unroll
      for (u32 j = 0; j < 4; ++j) rr[i][j] = gshfl(r[j], i);
  }
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    double const* in = L.src[t];
    double* out = L.dst[t];
    for (u32 k = 0; k < w; ++k) {
      double c[4] = {0, 0, 0, 0};
      double avg = 0;
      if (fit) {
        double b = row ? in[(size_t) ua[f + l] * w + k] : 0.0;
        double y[4];
#pragma unroll
        for (u32 i = 0; i < 4; ++i) {
          double p = q[i] * b;
          y[i] = 0;
#pragma unroll
          for (u32 j = 0; j < FIT_PTS; ++j) {
            double x = gshfl(p, j);
            if (j < npts) y[i] += x;
          }
        }
        // backsubst (qr.cpp:184-193)
#pragma unroll
        for (u32 ii = 0; ii < 4; ++ii) {
          u32 i = 3 - ii;
          c[i] = y[i];
#pragma unroll
          for (u32 j = i + 1; j < 4; ++j) c[i] -= rr[i][j] * c[j];
          c[i] /= rr[i][i];
        }
      } else {
        for (u32 j = f; j < e; ++j) avg += in[(size_t) ua[j] * w + k];
        avg /= (e - f);
      }
      for (u32 jj = l; jj < e - f; jj += FIT_PTS) {
        u32 el = ua[f + jj];
        if (eg[el] == eg[el + 1]) continue;
        u32 ne = eg[el] + nsame;
        double val;
        if (fit) {
          double const* xc = ncen + 3 * (size_t) ne;
          double d = 0;
          d += c[1] * xc[0];
          d += c[2] * xc[1];
          d += c[3] * xc[2];
          val = c[0] + d;
        } else {
          val = avg;
        }
        out[(size_t) ne * w + k] = val;
      }
    }
  }
}

// ---------------------------------------------------------------- swap
// swap_conserve_cavity / swap_fit_cavity: one thread per edge
__global__ __launch_bounds__(BLOCK) void swap_elem_fields(u32 ne, u32 const* __restrict__ ge,
                                                          u32 const* __restrict__ uo, u32 const* __restrict__ ua,
                                                          u32 nsame, double const* __restrict__ sizes,
                                                          FieldList L, int conserve) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= ne) return;
  u32 g0 = ge[i], g1 = ge[i + 1];
  if (g0 == g1) return;
  u32 f = uo[i], e = uo[i + 1];
  double cav = 0;
  if (conserve)
    for (u32 g = g0; g < g1; ++g) cav += sizes[nsame + g];
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    double const* in = L.src[t];
    double* out = L.dst[t];
    for (u32 k = 0; k < w; ++k) {
      double s = 0;
      if (conserve) {
        for (u32 j = f; j < e; ++j) s = s + in[(size_t) ua[j] * w + k];
        for (u32 g = g0; g < g1; ++g) out[(size_t) (nsame + g) * w + k] = s * (sizes[nsame + g] / cav);
      } else {
        for (u32 j = f; j < e; ++j) s += in[(size_t) ua[j] * w + k];
        s /= (e - f);
        for (u32 g = g0; g < g1; ++g) out[(size_t) (nsame + g) * w + k] = s;
      }
    }
  }
}

void run_copy_same(std::vector<FieldList> const& ls, u32 n, u32 const* soff) {
  for (auto const& L : ls)
    if (n) {
      hipLaunchKernelGGL(copy_same, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, soff, L);
      check_launch("transfer_copy_same");
    }
}

void check_ncomps(std::vector<Tag*> const& ts, char const* what) {
  for (Tag* t : ts) OHX_REQUIRE(t->ncomps <= MAX_NCOMPS, what);
}

}  // namespace

// refine_conserve then refine_fit (refine_common.cpp:110-113) for the
// elements: soff = survivor offsets, off = split offsets of the n old
// elements, children of split element i at go + off[i] * ppd + j
void refine_transfer(Mesh& m, Mesh& mo, u32 const* soff, u32 const* off, u32 ppd, u32 go, u32 nout) {
  u32 n = count(m, m.dim);
  static int const kinds[2] = {TR_CONSERVE, TR_POINTWISE};
  for (int kind : kinds) {
    auto ts = fields_of(m, kind);
    if (ts.empty()) continue;
    for (auto const& L : make_outputs(m, mo, ts, kind, nout))
      if (n) {
        hipLaunchKernelGGL(refine_elem_fields, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, soff, off, ppd, go, L,
                           kind == TR_CONSERVE ? 1 : 0);
        check_launch("refine_elem_fields");
      }
  }
}

// coarsen_conserve then coarsen_fit (coarsen_common.cpp:81-86);
// vgoff = offsets of the collapsing vertices, egoff / esoff = generated /
// survivor offsets of the old elements; mo already holds its vertices,
// their coordinates and its elements
void coarsen_transfer(Mesh& m, Mesh& mo, u32 const* vgoff, u32 ncol, u32 const* egoff, u32 const* esoff, u32 nsame,
                      u32 ngen) {
  u32 dim = m.dim, nv = count(m, 0), n = count(m, dim);
  auto cons = fields_of(m, TR_CONSERVE);
  auto fits = fields_of(m, TR_POINTWISE);
  if (cons.empty() && fits.empty()) return;
  Up const& up = ask_up(m, 0, dim);
  // the collapsing vertices, in vertex order
  Buf<u32> cols(ncol);
  {
    u32* cl = cols.get();
    parallel_for(nv, OHX_LAMBDA(u32 v) {
      if (vgoff[v] != vgoff[v + 1]) cl[vgoff[v]] = v;
    });
  }
  u32 nout = nsame + ngen;
  if (!cons.empty()) {
    check_ncomps(cons, "coarsen_conserve: more than 32 components");
    Buf<double> sizes = element_sizes(mo);
    auto ls = make_outputs(m, mo, cons, TR_CONSERVE, nout);
    run_copy_same(ls, n, esoff);
    for (auto const& L : ls)
      if (ncol) {
        hipLaunchKernelGGL(coarsen_conserve_cavity, dim3(nblocks(ncol)), dim3(BLOCK), 0, stream(), ncol, cols.get(),
                           up.offsets.get(), up.adj.get(), egoff, nsame, sizes.get(), L);
        check_launch("coarsen_conserve_cavity");
      }
  }
  if (fits.empty()) return;
  OHX_REQUIRE(dim == 3, "coarsen_fit: pointwise element fields need a 3D mesh");
  check_ncomps(fits, "coarsen_fit: more than 32 components");
  Tag* oc = find_tag(m, dim, "coordinates");
  if (!oc) oc = &add_centroids(m);
  Tag& nc = add_centroids(mo);
  auto ls = make_outputs(m, mo, fits, TR_POINTWISE, nout);
  run_copy_same(ls, n, esoff);
  for (auto const& L : ls)
    if (ncol) {
      hipLaunchKernelGGL(coarsen_fit_cavity, dim3(nblocks((size_t) ncol * FIT_PTS)), dim3(BLOCK), 0, stream(), ncol,
                         cols.get(), up.offsets.get(), up.adj.get(), egoff, nsame, oc->f.get(), nc.f.get(), L);
      check_launch("coarsen_fit_cavity");
    }
}

// swap_conserve then swap_fit (swap.cpp:63-66); egoff = products per edge
// (contiguous after the nsame kept elements), soff = survivor offsets
void swap_transfer(Mesh& m, Mesh& mo, u32 const* egoff, u32 const* soff, u32 nsame, u32 ngen) {
  u32 dim = m.dim, ne = count(m, 1), n = count(m, dim);
  auto cons = fields_of(m, TR_CONSERVE);
  auto fits = fields_of(m, TR_POINTWISE);
  if (cons.empty() && fits.empty()) return;
  Up const& up = ask_up(m, 1, dim);
  u32 nout = nsame + ngen;
  if (!cons.empty()) {
    check_ncomps(cons, "swap_conserve: more than 32 components");
    Buf<double> sizes = element_sizes(mo);
    auto ls = make_outputs(m, mo, cons, TR_CONSERVE, nout);
    run_copy_same(ls, n, soff);
    for (auto const& L : ls) {
      hipLaunchKernelGGL(swap_elem_fields, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, egoff, up.offsets.get(),
                         up.adj.get(), nsame, sizes.get(), L, 1);
      check_launch("swap_conserve");
    }
  }
  if (fits.empty()) return;
  OHX_REQUIRE(dim == 3, "swap_fit: pointwise element fields need a 3D mesh");
  auto ls = make_outputs(m, mo, fits, TR_POINTWISE, nout);
  run_copy_same(ls, n, soff);
  for (auto const& L : ls) {
    hipLaunchKernelGGL(swap_elem_fields, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, egoff, up.offsets.get(),
                       up.adj.get(), nsame, (double const*) nullptr, L, 0);
    check_launch("swap_fit");
  }
}

}  // namespace ohx
