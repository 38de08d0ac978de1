/* Test-infrastructure harness around the compiled reference (Omega_h v1,
 * serial or OpenMP loop build).  It is linked with the reference's own
 * objects by oracle/Makefile into oracle/_ref/libohv1_ref_*.so and is only
 * ever loaded by tests/, bench.py's cpu_baseline leg and the golden-fixture
 * script.  It exposes a flat C ABI (ref_*) over the reference's internal
 * seam: struct mesh + the pass functions of refine.hpp, coarsen.hpp,
 * swap.hpp, adapt.hpp, warp_to_limit.hpp and derive_model.hpp.
 *
 * Nothing in here re-implements reference algorithms except
 * ref_adapt_timed(), which repeats the control flow of mesh_adapt()
 * (adapt.cpp:22-94) with a clock around every pass so the CPU baseline
 * can be reported with the per-pass metric definition.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

#include "adapt.hpp"
#include "algebra.hpp"
#include "arrays.hpp"
#include "coarsen.hpp"
#include "derive_model.hpp"
#include "edge_swap.hpp"
#include "element_gradients.hpp"
#include "doubles.hpp"
#include "graph.hpp"
#include "indset.hpp"
#include "loop.hpp"
#include "mesh.hpp"
#include "mesh_diff.hpp"
#include "qr.hpp"
#include "quality.hpp"
#include "recover_by_volume.hpp"
#include "refine.hpp"
#include "size.hpp"
#include "size_from_hessian.hpp"
#include "swap.hpp"
#include "vtk_io.hpp"
#include "tables.hpp"
#include "tag.hpp"
#include "warp_to_limit.hpp"

using namespace omega_h;

#define M(p) reinterpret_cast<struct mesh*>(p)

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

extern "C" {

/* A serial (non-parallel) MESH_REDUCED mesh with vertices, elements and
 * coordinates, as new_box_mesh() builds (mesh.cpp:69-83) but from arrays. */
void* ref_new_mesh(unsigned dim, unsigned nverts, double const* coords,
    unsigned nelems, unsigned const* conn)
{
  struct mesh* m = new_mesh(dim, MESH_REDUCED, 0);
  mesh_set_ents(m, 0, nverts, 0);
  unsigned vpe = the_down_degrees[dim][0];
  mesh_set_ents(m, dim, nelems, array_to_device(conn, vpe * nelems));
  mesh_add_tag(m, 0, TAG_F64, "coordinates", 3,
      array_to_device(coords, 3 * nverts));
  return m;
}

void* ref_new_box_mesh(unsigned dim)
{
  return new_box_mesh(dim);
}

void ref_free(void* m) { free_mesh(M(m)); }

/* test_warp.cpp:105-110: classify with crease angle PI/4, then MESH_FULL */
void ref_derive_model_full(void* m)
{
  mesh_derive_model(M(m), PI / 4);
  mesh_set_rep(M(m), MESH_FULL);
}

void ref_derive_model(void* m, double crease_angle)
{
  mesh_derive_model(M(m), crease_angle);
}

void ref_set_full(void* m) { mesh_set_rep(M(m), MESH_FULL); }

unsigned ref_dim(void* m) { return mesh_dim(M(m)); }
unsigned ref_count(void* m, unsigned dim) { return mesh_count(M(m), dim); }
unsigned ref_has_dim(void* m, unsigned dim) { return mesh_has_dim(M(m), dim); }

void ref_get_down(void* m, unsigned h, unsigned l, unsigned* out)
{
  unsigned n = mesh_count(M(m), h) * the_down_degrees[h][l];
  unsigned const* d = mesh_ask_down(M(m), h, l);
  memcpy(out, d, sizeof(unsigned) * n);
}

unsigned ref_up_nuses(void* m, unsigned l, unsigned h)
{
  struct const_up* u = mesh_ask_up(M(m), l, h);
  return u->offsets[mesh_count(M(m), l)];
}

void ref_get_up(void* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj, unsigned* dirs)
{
  struct const_up* u = mesh_ask_up(M(m), l, h);
  unsigned n = mesh_count(M(m), l);
  unsigned nuses = u->offsets[n];
  memcpy(offs, u->offsets, sizeof(unsigned) * (n + 1));
  memcpy(adj, u->adj, sizeof(unsigned) * nuses);
  memcpy(dirs, u->directions, sizeof(unsigned) * nuses);
}

unsigned ref_star_nadj(void* m, unsigned l, unsigned h)
{
  struct const_graph* g = mesh_ask_star(M(m), l, h);
  return g->offsets[mesh_count(M(m), l)];
}

void ref_get_star(void* m, unsigned l, unsigned h, unsigned* offs, unsigned* adj)
{
  struct const_graph* g = mesh_ask_star(M(m), l, h);
  unsigned n = mesh_count(M(m), l);
  memcpy(offs, g->offsets, sizeof(unsigned) * (n + 1));
  memcpy(adj, g->adj, sizeof(unsigned) * g->offsets[n]);
}

void ref_get_dual(void* m, unsigned* out)
{
  unsigned d = mesh_dim(M(m));
  unsigned n = mesh_count(M(m), d) * the_down_degrees[d][d - 1];
  memcpy(out, mesh_ask_dual(M(m)), sizeof(unsigned) * n);
}

/* returns ncomps, or -1 if absent / wrong type */
int ref_get_tag(void* m, unsigned dim, char const* name, int want_f64, void* out)
{
  struct const_tag* t = mesh_find_tag(M(m), dim, name);
  if (!t)
    return -1;
  if ((t->type == TAG_F64) != (want_f64 != 0))
    return -1;
  if (out) {
    unsigned n = mesh_count(M(m), dim) * t->ncomps;
    memcpy(out, t->d.raw, n * tag_size(t->type));
  }
  return (int)t->ncomps;
}

unsigned ref_ntags(void* m, unsigned dim) { return mesh_count_tags(M(m), dim); }

char const* ref_tag_name(void* m, unsigned dim, unsigned i)
{
  return mesh_get_tag(M(m), dim, i)->name;
}

void ref_add_field(void* m, unsigned dim, char const* name, unsigned ncomps,
    double const* data)
{
  unsigned n = mesh_count(M(m), dim) * ncomps;
  mesh_add_tag(M(m), dim, TAG_F64, name, ncomps, array_to_device(data, n));
}

/* osh_new_field (omega_h.cpp:613-622): a field with a transfer type */
void ref_add_field_tt(void* m, unsigned dim, char const* name, unsigned ncomps,
    double const* data, int transfer)
{
  unsigned n = mesh_count(M(m), dim) * ncomps;
  add_tag2(mesh_tags(M(m), dim), TAG_F64, name, ncomps,
      static_cast<enum osh_transfer>(transfer), array_to_device(data, n));
}

void ref_free_field(void* m, unsigned dim, char const* name)
{
  mesh_free_tag(M(m), dim, name);
}

double ref_min_quality(void* m) { return mesh_min_quality(M(m)); }

void ref_qualities(void* m, double* out)
{
  unsigned n = mesh_count(M(m), mesh_dim(M(m)));
  double* q = mesh_qualities(M(m));
  memcpy(out, q, sizeof(double) * n);
  loop_free(q);
}

void ref_measure_edges(void* m, double* out)
{
  unsigned n = mesh_count(M(m), 1);
  double* s = mesh_measure_edges_for_adapt(M(m));
  memcpy(out, s, sizeof(double) * n);
  loop_free(s);
}

double ref_doubles_min(double const* a, unsigned n) { return doubles_min(a, n); }
double ref_doubles_max(double const* a, unsigned n) { return doubles_max(a, n); }

/* mesh_find_indset (indset.cpp:119-139) on the mesh's own star graph */
void ref_find_indset(void* m, unsigned ent_dim, unsigned const* candidates,
    double const* goodness, unsigned* out)
{
  unsigned n = mesh_count(M(m), ent_dim);
  unsigned* s = mesh_find_indset(M(m), ent_dim, candidates, goodness);
  memcpy(out, s, sizeof(unsigned) * n);
  loop_free(s);
}

unsigned ref_refine_by_size(void* m, double qual_floor)
{
  return refine_by_size(M(m), qual_floor);
}

void ref_uniformly_refine(void* m) { uniformly_refine(M(m)); }

unsigned ref_coarsen_by_size(void* m, double qual_floor, double size_floor)
{
  return coarsen_by_size(M(m), qual_floor, size_floor);
}

unsigned ref_coarsen_slivers(void* m, double qual_floor, unsigned nlayers)
{
  return coarsen_slivers(M(m), qual_floor, nlayers);
}

unsigned ref_swap_slivers(void* m, double good_qual, unsigned nlayers)
{
  return swap_slivers(M(m), good_qual, nlayers);
}

unsigned ref_adapt(void* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops)
{
  return mesh_adapt(M(m), size_floor, good_qual, nlayers, max_ops);
}

unsigned ref_warp_to_limit(void* m, double qual_floor)
{
  return mesh_warp_to_limit(M(m), qual_floor);
}

/* the reference's own swap tables and swap choice (edge_swap.cpp,
 * edge_swap_edges.cpp) and quality formulas (quality.hpp), for
 * known-answer fixtures */
unsigned ref_swap_nint_edges(unsigned ring_size)
{
  return swap_nint_edges[ring_size];
}

void ref_swap_int_edges(unsigned ring_size, unsigned code, unsigned* out)
{
  unsigned k = swap_nint_edges[ring_size];
  for (unsigned i = 0; i < 2 * k; ++i)
    out[i] = swap_int_edges[ring_size][code][i];
}

unsigned ref_choose_edge_swap(unsigned ring_size, double const* edge_x6,
    double const* ring_x, double* quality)
{
  double ex[2][3];
  double rx[MAX_EDGE_SWAP][3];
  memcpy(ex, edge_x6, sizeof(ex));
  memcpy(rx, ring_x, sizeof(double) * 3 * ring_size);
  struct swap_choice sc = choose_edge_swap(ring_size, ex, rx);
  *quality = sc.quality;
  return sc.code;
}

double ref_tet_quality(double const* x12)
{
  double x[4][3];
  memcpy(x, x12, sizeof(x));
  return tet_quality(x);
}

double ref_triangle_quality(double const* x9)
{
  double x[3][3];
  memcpy(x, x9, sizeof(x));
  return triangle_quality(x);
}

double ref_triangle_xy_quality(double const* x9)
{
  double x[3][3];
  memcpy(x, x9, sizeof(x));
  return triangle_xy_quality(x);
}

/* Same control flow as mesh_adapt (adapt.cpp:82-94) with satisfy_size
 * (47-56), satisfy_shape (58-80) and adapt_summary (22-34, without its
 * printf), with CLOCK_MONOTONIC around every pass function call.
 * stats[0] = number of passes that changed the mesh,
 * stats[1] = sum of input element counts over those passes,
 * stats[2] = seconds spent in those passes,
 * stats[3] = seconds of the whole call.
 * Returns -1 instead of aborting when max_ops is exceeded (with the stats of
 * the passes run until then). */
static void summary_quiet(struct mesh* m)
{
  (void) mesh_min_quality(m);
  unsigned nedges = mesh_count(m, 1);
  double* s = mesh_measure_edges_for_adapt(m);
  (void) doubles_min(s, nedges);
  (void) doubles_max(s, nedges);
  loop_free(s);
}

int ref_adapt_timed(void* vm, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops, double* stats)
{
  struct mesh* m = M(vm);
  double t_begin = now_s();
  unsigned ops = 0;
  double npass = 0, nelems = 0, tpass = 0;
  summary_quiet(m);
#define TIMED_PASS(call, ok) do { \
    double tin = (double) mesh_count(m, mesh_dim(m)); \
    double t0 = now_s(); \
    ok = (call); \
    double t1 = now_s(); \
    if (ok) { npass += 1; nelems += tin; tpass += t1 - t0; } \
  } while (0)
#define COUNT_OP() do { \
    if (ops > max_ops) { \
      stats[0] = npass; stats[1] = nelems; stats[2] = tpass; stats[3] = now_s() - t_begin; \
      return -1; } \
    ++ops; summary_quiet(m); } while (0)
  {
    double qual_floor = mesh_min_quality(m);
    if (good_qual < qual_floor)
      qual_floor = good_qual;
    unsigned ok;
    while (1) { TIMED_PASS(refine_by_size(m, qual_floor), ok); if (!ok) break; COUNT_OP(); }
    while (1) { TIMED_PASS(coarsen_by_size(m, qual_floor, size_floor), ok); if (!ok) break; COUNT_OP(); }
  }
  while (1) {
    double prev_qual = mesh_min_quality(m);
    if (prev_qual >= good_qual)
      break;
    unsigned ok = 0;
    if (mesh_dim(m) == 3) {
      TIMED_PASS(swap_slivers(m, good_qual, nlayers), ok);
      if (ok) { COUNT_OP(); continue; }
    }
    TIMED_PASS(coarsen_slivers(m, good_qual, nlayers), ok);
    if (ok) { COUNT_OP(); continue; }
    break;
  }
#undef TIMED_PASS
#undef COUNT_OP
  stats[0] = npass;
  stats[1] = nelems;
  stats[2] = tpass;
  stats[3] = now_s() - t_begin;
  return ops > 0;
}


/* read_mesh_vtk / write_mesh_vtk (vtk_io.cpp:787-830), for the file-format
 * parity tests */
void* ref_read_vtk(char const* path)
{
  return read_mesh_vtk(path);
}

void ref_write_vtk(void* m, char const* path)
{
  write_mesh_vtk(static_cast<struct mesh*>(m), path);
}

/* the size-field producers of the warp loop (test_warp.cpp:70-83) and
 * the identity size (size.cpp:90-135) */
void ref_element_gradients(void* m, char const* name) { mesh_element_gradients(M(m), name); }
void ref_recover_by_volume(void* m, char const* name) { mesh_recover_by_volume(M(m), name); }
void ref_size_from_hessian(void* m, char const* hess_name, double const* weights,
    double min_h, double max_h)
{
  mesh_size_from_hessian(M(m), hess_name, weights, min_h, max_h);
}
void ref_identity_size(void* m, char const* name) { mesh_identity_size_field(M(m), name); }

/* qr_decomp2 + qr_solve2 (qr.cpp:151-200) on an npts x 4 row-major matrix:
 * the rank found, and the solution when it is 4 */
unsigned ref_qr_fit(unsigned npts, double const* a, double const* b, double* x)
{
  double aa[MAX_PTS][4], q[MAX_PTS][MAX_PTS], r[MAX_PTS][4], bb[MAX_PTS];
  for (unsigned i = 0; i < npts; ++i) {
    for (unsigned j = 0; j < 4; ++j)
      aa[i][j] = a[i * 4 + j];
    bb[i] = b[i];
  }
  unsigned rank = qr_decomp2(aa, q, r, npts);
  if (rank == 4)
    qr_solve2(q, r, bb, npts, x);
  return rank;
}

/* mesh_diff (mesh_diff.cpp:141-178), stdout flushed so a caller sees the
 * messages in order */
unsigned ref_mesh_diff(void* a, void* b, double tol, double floor, unsigned superset)
{
  unsigned r = mesh_diff(M(a), M(b), tol, floor, superset);
  fflush(stdout);
  return r;
}
}
