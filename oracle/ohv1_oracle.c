/* CPU oracle for the Omega_h v1 adapt-pass hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, bench.py (cpu_baseline) and
 * __graft_entry__.smoke() as the checker; never by the product.
 *
 * A from-scratch serial C99 restatement of the reference's behaviour: FP64,
 * isotropic per-vertex "adapt_size", the serial loop build (including the
 * serial min/max that skip element 0, doubles.cpp:100-118).  Compile with
 * -ffp-contract=off so every floating-point operation rounds like the
 * reference's x86-64 SSE2 build.  Function comments cite the reference
 * file:line each one restates.
 */
#define _POSIX_C_SOURCE 199309L
#include "ohv1_oracle.h"

#include <assert.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned u32;

#define INVALID (~0u)
#define MAXUP 128
#define REF_PI 3.14159265359 /* algebra.hpp:8 */

static int g_verbose = 0;
static unsigned indset_max_rounds = 100;
/* round limit of the independent set (indset.cpp:33: 100); raising it is an
 * extension for meshes whose ties form long chains */
void orc_set_indset_max_rounds(unsigned n)
{
  indset_max_rounds = n ? n : 100;
}

void orc_set_verbose(int v) { g_verbose = v; }

static void die(char const* what)
{
  fprintf(stderr, "oracle: %s\n", what);
  abort();
}

static void* xmalloc(size_t n)
{
  void* p = malloc(n ? n : 1);
  if (!p)
    die("out of memory");
  return p;
}
#define NEW(T, n) ((T*) xmalloc(sizeof(T) * (size_t)(n)))

void orc_free_buf(void* p) { free(p); }

/* ---------------------------------------------------------------- tables
 * Local orderings of simplex sub-entities (tables.cpp:77-193).  These must
 * be identical to the reference's since they define every output order.
 * canon(h,l,w)[j] lists the w-dimensional sub-entities (by their local index
 * in h) of the j-th l-dimensional sub-entity of an h-simplex. */
static u32 const down_deg[4][4] = {
  {1, 0, 0, 0}, {2, 1, 0, 0}, {3, 3, 1, 0}, {4, 6, 4, 1}};

static u32 const c_f_e_v[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static u32 const c_r_e_v[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
static u32 const c_r_f_v[4][3] = {{0, 1, 2}, {1, 0, 3}, {2, 1, 3}, {0, 2, 3}};
static u32 const c_r_f_e[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 4, 5}, {2, 5, 3}};
static u32 const iota6[6] = {0, 1, 2, 3, 4, 5};

/* returns row j of the (h,l,w) table.  Rows of the form (h,l,l) are the
 * identity {j}; rows (h,h,w) list all w-entities of the simplex in order. */
static u32 const* canon(u32 h, u32 l, u32 w, u32 j)
{
  if (w == l)
    return &iota6[(h == l) ? 0 : j];
  if (h == l)
    return iota6;
  switch (h * 100 + l * 10 + w) {
    case 210: return c_f_e_v[j];
    case 310: return c_r_e_v[j];
    case 320: return c_r_f_v[j];
    case 321: return c_r_f_e[j];
  }
  die("bad canonical order request");
  return 0;
}
#define CANON canon

/* the_opposite_orders (tables.cpp:181-193): local index of the
 * sub-entity opposite to the l-dimensional one */
static u32 const o_e_v[2] = {1, 0};
static u32 const o_f_v[3] = {1, 2, 0};
static u32 const o_f_e[3] = {2, 0, 1};
static u32 const o_r_v[4] = {2, 3, 1, 0};
static u32 const o_r_e[6] = {5, 3, 4, 1, 2, 0};
static u32 const o_r_f[4] = {3, 2, 0, 1};
static u32 const* opp(u32 h, u32 l)
{
  switch (h * 10 + l) {
    case 10: return o_e_v;
    case 20: return o_f_v;
    case 21: return o_f_e;
    case 30: return o_r_v;
    case 31: return o_r_e;
    case 32: return o_r_f;
  }
  die("bad opposite order request");
  return 0;
}

/* ---------------------------------------------------------- generator */
void orc_box(unsigned dim, unsigned n, unsigned* p_nverts, double** p_coords,
    unsigned* p_nelems, unsigned** p_conn)
{
  /* corner numbering and per-cube split of box_2d/3d (tables.cpp:15-47),
   * repeated over an n^dim lattice; vertex (i,j,k) has index
   * i + (n+1)*(j + (n+1)*k) and coordinates (i/n, j/n, k/n). */
  static u32 const c2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  static u32 const t2[2][3] = {{0, 1, 2}, {2, 3, 0}};
  static u32 const c3[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  static u32 const t3[6][4] = {{0, 1, 2, 6}, {2, 3, 0, 6}, {0, 3, 7, 6},
    {7, 4, 0, 6}, {0, 4, 5, 6}, {5, 1, 0, 6}};
  u32 np = n + 1;
  u32 nverts = (dim == 3) ? np * np * np : np * np;
  double* x = NEW(double, 3 * nverts);
  for (u32 v = 0; v < nverts; ++v) {
    u32 i = v % np, j = (v / np) % np, k = v / (np * np);
    x[v * 3 + 0] = (double) i / (double) n;
    x[v * 3 + 1] = (double) j / (double) n;
    x[v * 3 + 2] = (dim == 3) ? (double) k / (double) n : 0.0;
  }
  u32 ncubes = (dim == 3) ? n * n * n : n * n;
  u32 per = (dim == 3) ? 6 : 2;
  u32 vpe = dim + 1;
  u32 nelems = ncubes * per;
  u32* conn = NEW(u32, nelems * vpe);
  for (u32 c = 0; c < ncubes; ++c) {
    u32 ci = c % n, cj = (c / n) % n, ck = c / (n * n);
    u32 vid[8];
    if (dim == 3) {
      for (u32 a = 0; a < 8; ++a)
        vid[a] = (ci + c3[a][0]) + np * ((cj + c3[a][1]) + np * (ck + c3[a][2]));
      for (u32 t = 0; t < 6; ++t)
        for (u32 a = 0; a < 4; ++a)
          conn[(c * 6 + t) * 4 + a] = vid[t3[t][a]];
    } else {
      for (u32 a = 0; a < 4; ++a)
        vid[a] = (ci + c2[a][0]) + np * (cj + c2[a][1]);
      for (u32 t = 0; t < 2; ++t)
        for (u32 a = 0; a < 3; ++a)
          conn[(c * 2 + t) * 3 + a] = vid[t2[t][a]];
    }
  }
  *p_nverts = nverts;
  *p_coords = x;
  *p_nelems = nelems;
  *p_conn = conn;
}

/* ------------------------------------------------------- array helpers
 * ints.cpp:25-246 (serial branch), arrays.cpp:132-233 */
static u32* u32_exscan(u32 const* a, u32 n) /* ints.cpp:25-36 */
{
  u32* o = NEW(u32, n + 1);
  u32 s = 0;
  o[0] = 0;
  for (u32 i = 0; i < n; ++i) {
    s += a[i];
    o[i + 1] = s;
  }
  return o;
}

static u32 u32_max(u32 const* a, u32 n) /* ints.cpp:168-175 */
{
  u32 m = 0;
  for (u32 i = 0; i < n; ++i)
    if (a[i] > m)
      m = a[i];
  return m;
}

static u32 u32_sum(u32 const* a, u32 n) /* ints.cpp:182-188 */
{
  u32 s = 0;
  for (u32 i = 0; i < n; ++i)
    s += a[i];
  return s;
}

/* uints_negate_offsets (ints.cpp:224-232): offsets of the entities whose
 * own offset range is empty */
static u32* u32_negate_offsets(u32 const* offs, u32 n)
{
  u32* neg = NEW(u32, n);
  for (u32 i = 0; i < n; ++i)
    neg[i] = !(offs[i + 1] - offs[i]);
  u32* o = u32_exscan(neg, n);
  free(neg);
  return o;
}

static u32* u32_scale(u32 const* a, u32 n, u32 s) /* ints.cpp:234-244 */
{
  u32* o = NEW(u32, n);
  for (u32 i = 0; i < n; ++i)
    o[i] = s * a[i];
  return o;
}

static u32* u32_filled(u32 n, u32 v)
{
  u32* o = NEW(u32, n);
  for (u32 i = 0; i < n; ++i)
    o[i] = v;
  return o;
}

static u32* u32_copy(u32 const* a, u32 n)
{
  u32* o = NEW(u32, n);
  memcpy(o, a, sizeof(u32) * n);
  return o;
}

static double* f64_copy(double const* a, u32 n)
{
  double* o = NEW(double, n);
  memcpy(o, a, sizeof(double) * n);
  return o;
}

/* expand_array (arrays.cpp:187-213): entity i's width values are copied
 * into every slot of offsets[i]..offsets[i+1] */
static void* expand(void const* a, size_t esz, u32 const* offs, u32 n, u32 width)
{
  u32 nout = offs[n];
  char* out = (char*) xmalloc(esz * (size_t) nout * width);
  char const* in = (char const*) a;
  for (u32 i = 0; i < n; ++i)
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      memcpy(out + (size_t) j * width * esz, in + (size_t) i * width * esz,
          esz * width);
  return out;
}

static void* concat(void const* a, void const* b, size_t esz, u32 na, u32 nb,
    u32 width) /* arrays.cpp:215-224 */
{
  char* out = (char*) xmalloc(esz * ((size_t) na + nb) * width);
  memcpy(out, a, esz * (size_t) na * width);
  memcpy(out + esz * (size_t) na * width, b, esz * (size_t) nb * width);
  return out;
}

/* serial doubles_min / doubles_max: loops start at 1 (doubles.cpp:100-118) */
double orc_doubles_min(double const* a, unsigned n)
{
  double m = DBL_MAX;
  for (u32 i = 1; i < n; ++i)
    if (a[i] < m)
      m = a[i];
  return m;
}

double orc_doubles_max(double const* a, unsigned n)
{
  double m = -DBL_MAX;
  for (u32 i = 1; i < n; ++i)
    if (a[i] > m)
      m = a[i];
  return m;
}

/* ------------------------------------------------------------- tags
 * tag.cpp:12-120: an ordered list of named typed arrays per dimension */
enum { T_U32 = 1, T_F64 = 3 };

typedef struct {
  char name[64];
  int type;
  u32 ncomps;
  int transfer; /* enum osh_transfer: 0 not, 1 pointwise, 2 conserve */
  void* data;
} otag;

typedef struct {
  u32 n;
  otag* at[32];
} otags;

static otag* tag_find(otags* ts, char const* name)
{
  for (u32 i = 0; i < ts->n; ++i)
    if (!strcmp(ts->at[i]->name, name))
      return ts->at[i];
  return 0;
}

/* add_tag2 (tag.cpp:64-77) */
static void tag_add2(otags* ts, int type, char const* name, u32 ncomps, int transfer,
    void* data)
{
  if (tag_find(ts, name))
    die("tag added twice");
  if (ts->n == 32)
    die("too many tags");
  otag* t = NEW(otag, 1);
  snprintf(t->name, sizeof(t->name), "%s", name);
  t->type = type;
  t->ncomps = ncomps;
  t->transfer = transfer;
  t->data = data;
  ts->at[ts->n++] = t;
}

/* add_tag (tag.cpp:58-62): transfer type OSH_TRANSFER_NOT */
static void tag_add(otags* ts, int type, char const* name, u32 ncomps, void* data)
{
  tag_add2(ts, type, name, ncomps, 0, data);
}

static void tag_remove(otags* ts, char const* name)
{
  for (u32 i = 0; i < ts->n; ++i)
    if (!strcmp(ts->at[i]->name, name)) {
      free(ts->at[i]->data);
      free(ts->at[i]);
      for (u32 j = i; j + 1 < ts->n; ++j)
        ts->at[j] = ts->at[j + 1];
      ts->n--;
      return;
    }
  die("removing a missing tag");
}

static void tags_free(otags* ts)
{
  for (u32 i = 0; i < ts->n; ++i) {
    free(ts->at[i]->data);
    free(ts->at[i]);
  }
  ts->n = 0;
}

static size_t tag_esz(int type) { return type == T_F64 ? sizeof(double) : sizeof(u32); }

/* ------------------------------------------------------------- mesh
 * mesh.cpp:21-334: counts, lazily derived adjacency caches, tags */
typedef struct {
  u32* offsets;
  u32* adj;
  u32* dirs;
} oup;

typedef struct {
  u32* offsets;
  u32* adj;
} ograph;

struct orc_mesh {
  u32 dim;
  int full;
  u32 counts[4];
  u32* down[4][4];
  oup* up[4][4];
  ograph* star[4][4];
  u32* dual;
  otags tags[4];
};

static orc_mesh* mesh_new(u32 dim, int full)
{
  orc_mesh* m = NEW(orc_mesh, 1);
  memset(m, 0, sizeof(*m));
  m->dim = dim;
  m->full = full;
  return m;
}

static void mesh_free_contents(orc_mesh* m)
{
  for (u32 h = 0; h < 4; ++h)
    for (u32 l = 0; l < 4; ++l) {
      free(m->down[h][l]);
      if (m->up[l][h]) {
        free(m->up[l][h]->offsets);
        free(m->up[l][h]->adj);
        free(m->up[l][h]->dirs);
        free(m->up[l][h]);
      }
      if (m->star[l][h]) {
        free(m->star[l][h]->offsets);
        free(m->star[l][h]->adj);
        free(m->star[l][h]);
      }
    }
  free(m->dual);
  for (u32 d = 0; d < 4; ++d)
    tags_free(&m->tags[d]);
}

void orc_free(orc_mesh* m)
{
  mesh_free_contents(m);
  free(m);
}

/* overwrite_mesh (mesh.cpp:329-334) */
static void mesh_overwrite(orc_mesh* m, orc_mesh* with)
{
  mesh_free_contents(m);
  *m = *with;
  free(with);
}

static void mesh_set_ents(orc_mesh* m, u32 dim, u32 n, u32* verts)
{
  m->counts[dim] = n;
  if (m->down[dim][0])
    die("entities set twice");
  m->down[dim][0] = verts;
}

static int mesh_has_dim(orc_mesh* m, u32 d) { return d == 0 || m->down[d][0] != 0; }

static u32 const* ask_down(orc_mesh* m, u32 h, u32 l);
static oup const* ask_up(orc_mesh* m, u32 l, u32 h);
static ograph const* ask_star(orc_mesh* m, u32 l, u32 h);
static u32 const* ask_dual(orc_mesh* m);

static u32 mesh_count(orc_mesh* m, u32 d) /* mesh.cpp:89-99 */
{
  if (d > m->dim)
    return 0;
  if (d)
    ask_down(m, d, 0);
  return m->counts[d];
}

static otag* find_tag(orc_mesh* m, u32 d, char const* name)
{
  return tag_find(&m->tags[d], name);
}

static double* coords_of(orc_mesh* m)
{
  otag* t = find_tag(m, 0, "coordinates");
  if (!t)
    die("no coordinates");
  return (double*) t->data;
}

/* ---------------------------------------------------------- algebra
 * algebra.hpp, size.hpp, quality.hpp -- identical operation order */
static double vsqdist(double const* a, double const* b)
{
  double s = 0;
  for (u32 i = 0; i < 3; ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

static void cross(double const* a, double const* b, double* c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(double const* a, double const* b)
{
  double d = 0;
  for (u32 i = 0; i < 3; ++i)
    d += a[i] * b[i];
  return d;
}

static void sub3(double const* a, double const* b, double* c)
{
  for (u32 i = 0; i < 3; ++i)
    c[i] = a[i] - b[i];
}

static double tri_area(double const (*x)[3]) /* size.hpp:14-23 */
{
  double v0[3], v1[3], c[3];
  sub3(x[1], x[0], v0);
  sub3(x[2], x[0], v1);
  cross(v0, v1, c);
  return sqrt(dot3(c, c)) / 2.0;
}

static double tri_z_area(double const (*x)[3]) /* size.hpp:25-34 */
{
  double v0[3], v1[3], c[3];
  sub3(x[1], x[0], v0);
  sub3(x[2], x[0], v1);
  cross(v0, v1, c);
  return c[2] / 2.0;
}

static double tet_volume(double const (*x)[3]) /* size.hpp:36-46 */
{
  double v0[3], v1[3], v2[3], c[3];
  sub3(x[1], x[0], v0);
  sub3(x[2], x[0], v1);
  sub3(x[3], x[0], v2);
  cross(v0, v1, c);
  return dot3(c, v2) / 6.0;
}

/* PERFECT_TRIANGLE_QUALITY, PERFECT_TET_QUALITY (quality.hpp:15-18) */
static double perfect_tri(void) { return sqrt(3.0) / 4.0; }
static double perfect_tet(void)
{
  double s = sqrt(perfect_tri());
  return (sqrt(2.0) / 12.0) / (s * s * s);
}

static double tri_quality_gen(double const (*x)[3], int z) /* quality.hpp:34-66 */
{
  double sum_lsq = 0;
  for (u32 i = 0; i < 3; ++i)
    sum_lsq += vsqdist(x[i], x[(i + 1) % 3]);
  double lrms = sqrt(sum_lsq / 3);
  double a = z ? tri_z_area(x) : tri_area(x);
  double q = a / (lrms * lrms);
  return q / perfect_tri();
}

static double tet_quality(double const (*x)[3]) /* quality.hpp:68-89 */
{
  double sum_asq = 0;
  for (u32 i = 0; i < 4; ++i) {
    double tx[3][3];
    for (u32 j = 0; j < 3; ++j)
      memcpy(tx[j], x[c_r_f_v[i][j]], sizeof(double) * 3);
    double a = tri_area((double const (*)[3]) tx);
    sum_asq += a * a;
  }
  double arms = sqrt(sum_asq / 4.);
  double v = tet_volume(x);
  double r = sqrt(arms);
  double q = v / (r * r * r);
  return q / perfect_tet();
}

/* entity_quality uses the unsigned triangle area, element_quality the
 * signed z-area (quality.hpp:91-109) */
static double entity_quality(u32 dim, double const (*x)[3])
{
  return dim == 3 ? tet_quality(x) : tri_quality_gen(x, 0);
}
static double element_quality(u32 dim, double const (*x)[3])
{
  return dim == 3 ? tet_quality(x) : tri_quality_gen(x, 1);
}

/* ---------------------------------------------- anisotropic extension
 * Not in the reference (which has only the scalar "adapt_size").  A vertex
 * field "metric" holds an SPD 3x3 tensor (m00, m01, m02, m11, m12, m22); an
 * element's metric is its vertices' average (summed in vertex order, then
 * divided); its quality is the reference's formula on the edge vectors
 * mapped by the upper Cholesky factor U (U^T U = M); an edge's metric length
 * is sqrt(e^T M e) with M the average of its two vertices'.  Same operation
 * order as the device code. */
static double const* metric_of(orc_mesh* m)
{
  otag* t = find_tag(m, 0, "metric");
  if (!t)
    return 0;
  if (t->type != T_F64 || t->ncomps != 6)
    die("metric field must have 6 FP64 components");
  return (double const*) t->data;
}

static void sym_average(double const (*mv)[6], u32 n, double* out)
{
  for (u32 k = 0; k < 6; ++k) {
    double a = 0;
    for (u32 i = 0; i < n; ++i)
      a += mv[i][k];
    a /= n;
    out[k] = a;
  }
}

static void chol_map(double const* mm, double const* e, double* y)
{
  double u00 = sqrt(mm[0]);
  double u01 = mm[1] / u00;
  double u02 = mm[2] / u00;
  double u11 = sqrt(mm[3] - u01 * u01);
  double u12 = (mm[4] - u01 * u02) / u11;
  double u22 = sqrt((mm[5] - u02 * u02) - u12 * u12);
  y[0] = (u00 * e[0] + u01 * e[1]) + u02 * e[2];
  y[1] = u11 * e[1] + u12 * e[2];
  y[2] = u22 * e[2];
}

static double metric_quality(u32 dim, double const (*x)[3], double const (*mv)[6], int signed_z)
{
  double me[6];
  sym_average(mv, dim + 1, me);
  double p[4][3] = {{0, 0, 0}};
  for (u32 i = 1; i <= dim; ++i) {
    double e[3];
    sub3(x[i], x[0], e);
    chol_map(me, e, p[i]);
  }
  if (dim == 3)
    return tet_quality((double const (*)[3]) p);
  return tri_quality_gen((double const (*)[3]) p, signed_z);
}

static double metric_len2(double const* a, double const* b, double const* e)
{
  double mm[6];
  for (u32 k = 0; k < 6; ++k) {
    double s = 0;
    s += a[k];
    s += b[k];
    s /= 2;
    mm[k] = s;
  }
  double s = 0;
  s += e[0] * ((mm[0] * e[0] + mm[1] * e[1]) + mm[2] * e[2]);
  s += e[1] * ((mm[1] * e[0] + mm[3] * e[1]) + mm[4] * e[2]);
  s += e[2] * ((mm[2] * e[0] + mm[4] * e[1]) + mm[5] * e[2]);
  return s;
}

/* quality of an element given its corner coordinates and, if mt, corner
 * metrics */
static double corner_quality(u32 dim, double const (*x)[3], double const (*mv)[6],
    double const* mt, int signed_z)
{
  if (mt)
    return metric_quality(dim, x, mv, signed_z);
  return signed_z ? element_quality(dim, x) : entity_quality(dim, x);
}

double orc_tet_quality(double const* x12) { return tet_quality((double const (*)[3]) x12); }
double orc_triangle_quality(double const* x9) { return tri_quality_gen((double const (*)[3]) x9, 0); }
double orc_triangle_xy_quality(double const* x9) { return tri_quality_gen((double const (*)[3]) x9, 1); }

/* element_qualities (quality.cpp:13-45) */
static double* element_qualities(u32 dim, u32 n, u32 const* ev, double const* x,
    double const* mt)
{
  double* q = NEW(double, n);
  u32 vpe = dim + 1;
  for (u32 i = 0; i < n; ++i) {
    double ex[4][3], em[4][6];
    for (u32 j = 0; j < vpe; ++j) {
      memcpy(ex[j], x + 3 * ev[i * vpe + j], sizeof(double) * 3);
      if (mt)
        memcpy(em[j], mt + 6 * ev[i * vpe + j], sizeof(double) * 6);
    }
    q[i] = corner_quality(dim, (double const (*)[3]) ex, (double const (*)[6]) em, mt, 1);
  }
  return q;
}

static double* mesh_qualities(orc_mesh* m)
{
  return element_qualities(m->dim, mesh_count(m, m->dim), ask_down(m, m->dim, 0),
      coords_of(m), metric_of(m));
}

static double mesh_min_quality(orc_mesh* m) /* quality.cpp:67-73 */
{
  u32 n = mesh_count(m, m->dim);
  double* q = mesh_qualities(m);
  double mq = orc_doubles_min(q, n);
  free(q);
  return mq;
}

/* measure_edge (size.cpp:57-88): |x1-x0| / ((h0+h1)/2) */
static double* measure_edges(orc_mesh* m)
{
  u32 ne = mesh_count(m, 1);
  u32 const* ev = ask_down(m, 1, 0);
  double const* x = coords_of(m);
  double const* mt = metric_of(m);
  double const* h = 0;
  if (!mt) {
    otag* st = find_tag(m, 0, "adapt_size");
    if (!st)
      die("no adapt_size and no metric");
    h = (double const*) st->data;
  }
  double* out = NEW(double, ne);
  for (u32 i = 0; i < ne; ++i) {
    u32 a = ev[i * 2 + 0], b = ev[i * 2 + 1];
    if (mt) {
      double e[3];
      sub3(x + 3 * b, x + 3 * a, e);
      out[i] = sqrt(metric_len2(mt + 6 * a, mt + 6 * b, e));
      continue;
    }
    double len = sqrt(vsqdist(x + 3 * b, x + 3 * a));
    double want = (h[a] + h[b]) / 2;
    out[i] = len / want;
  }
  return out;
}

/* ------------------------------------------------- topology derivation */

/* invert_map (invert_map.cpp:72-136, host branch): bucket i of the output
 * lists the positions j with in[j] == i.  The reference fills with atomics
 * and then sorts each bucket ascending, so the result is the stable
 * counting sort of positions by value. */
void orc_invert_map(unsigned nin, unsigned const* in, unsigned nout,
    unsigned* out, unsigned* offsets)
{
  u32* counts = u32_filled(nout, 0);
  for (u32 i = 0; i < nin; ++i) {
    if (in[i] >= nout)
      die("invert_map: value out of range");
    counts[in[i]]++;
  }
  offsets[0] = 0;
  for (u32 i = 0; i < nout; ++i)
    offsets[i + 1] = offsets[i] + counts[i];
  for (u32 i = 0; i < nout; ++i)
    counts[i] = 0;
  for (u32 i = 0; i < nin; ++i) {
    u32 d = in[i];
    out[offsets[d] + counts[d]++] = i;
  }
  free(counts);
}

/* up_from_down (up_from_down.cpp:9-47): use j of a low entity is the pair
 * (high, direction) = (j / deg, j % deg) */
static oup* up_from_down(u32 hd, u32 ld, u32 nhighs, u32 nlows, u32 const* lows_of_highs)
{
  u32 deg = down_deg[hd][ld];
  u32 nuses = nhighs * deg;
  oup* u = NEW(oup, 1);
  u->offsets = NEW(u32, nlows + 1);
  u->adj = NEW(u32, nuses);
  u->dirs = NEW(u32, nuses);
  orc_invert_map(nuses, lows_of_highs, nlows, u->adj, u->offsets);
  for (u32 i = 0; i < nuses; ++i) {
    u32 both = u->adj[i];
    u->adj[i] = both / deg;
    u->dirs[i] = both % deg;
  }
  return u;
}

/* find_edge_up / find_tri_up (find_by_verts.hpp:14-76): scan the low
 * entities adjacent to the first wanted vertex */
static u32 find_by_verts(u32 ld, u32 const* want, u32 const* verts_of_lows,
    oup const* lv)
{
  u32 v0 = want[0];
  for (u32 i = lv->offsets[v0]; i < lv->offsets[v0 + 1]; ++i) {
    u32 e = lv->adj[i];
    u32 dir = lv->dirs[i];
    if (ld == 1) {
      if (verts_of_lows[e * 2 + (1 - dir)] == want[1])
        return e;
    } else {
      u32 const* t = verts_of_lows + e * 3;
      u32 a = t[(dir + 1) % 3], b = t[(dir + 2) % 3];
      if ((a == want[1] && b == want[2]) || (a == want[2] && b == want[1]))
        return e;
    }
  }
  die("find_by_verts: not found");
  return INVALID;
}

/* reflect_down (reflect_down.cpp:13-81) */
static u32* reflect_down(orc_mesh* m, u32 hd, u32 ld)
{
  u32 nh = mesh_count(m, hd);
  u32 const* vh = ask_down(m, hd, 0);
  oup const* lv = ask_up(m, 0, ld);
  u32 const* vl = ask_down(m, ld, 0);
  u32 vph = down_deg[hd][0], lph = down_deg[hd][ld], vpl = down_deg[ld][0];
  u32* out = NEW(u32, nh * lph);
  for (u32 i = 0; i < nh; ++i)
    for (u32 j = 0; j < lph; ++j) {
      u32 const* lc = CANON(hd, ld, 0, j);
      u32 want[3];
      for (u32 k = 0; k < vpl; ++k)
        want[k] = vh[i * vph + lc[k]];
      out[i * lph + j] = find_by_verts(ld, want, vl, lv);
    }
  return out;
}

static u32 add_unique(u32* a, u32 n, u32 e) /* ints.hpp:19-35 */
{
  for (u32 i = 0; i < n; ++i)
    if (a[i] == e)
      return n;
  a[n] = e;
  return n + 1;
}

/* get_star_general (star.cpp:13-111): for each low, lows of its highs in
 * up-adjacency order, first occurrence kept */
static ograph* star_general(orc_mesh* m, u32 ld, u32 hd)
{
  u32 nl = mesh_count(m, ld);
  oup const* u = ask_up(m, ld, hd);
  u32 const* lh = ask_down(m, hd, ld);
  u32 lph = down_deg[hd][ld];
  u32* buf = NEW(u32, MAXUP * 6);
  u32* deg = NEW(u32, nl);
  for (int pass = 0; pass < 2; ++pass) {
    ograph* g = 0;
    if (pass == 1) {
      g = NEW(ograph, 1);
      g->offsets = u32_exscan(deg, nl);
      g->adj = NEW(u32, g->offsets[nl]);
    }
    for (u32 i = 0; i < nl; ++i) {
      u32 size = 0;
      u32* dst = pass ? g->adj + g->offsets[i] : buf;
      for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
        u32 h = u->adj[j];
        for (u32 k = 0; k < lph; ++k) {
          u32 sl = lh[h * lph + k];
          if (sl == i)
            continue;
          size = add_unique(dst, size, sl);
        }
      }
      if (!pass)
        deg[i] = size;
    }
    if (pass == 1) {
      free(buf);
      free(deg);
      return g;
    }
  }
  return 0;
}

/* get_vertex_edge_star (star.cpp:113-153) */
static ograph* star_vertex_edge(orc_mesh* m)
{
  u32 nv = mesh_count(m, 0);
  u32 const* ev = ask_down(m, 1, 0);
  oup const* u = ask_up(m, 0, 1);
  ograph* g = NEW(ograph, 1);
  g->offsets = u32_copy(u->offsets, nv + 1);
  g->adj = NEW(u32, u->offsets[nv]);
  for (u32 i = 0; i < nv; ++i)
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j)
      g->adj[j] = ev[u->adj[j] * 2 + (1 - u->dirs[j])];
  return g;
}

/* get_edge_triangle_star (star.cpp:155-200) */
static ograph* star_edge_tri(orc_mesh* m)
{
  u32 ne = mesh_count(m, 1);
  u32 const* et = ask_down(m, 2, 1);
  oup const* u = ask_up(m, 1, 2);
  ograph* g = NEW(ograph, 1);
  g->offsets = NEW(u32, ne + 1);
  for (u32 i = 0; i <= ne; ++i)
    g->offsets[i] = u->offsets[i] * 2;
  g->adj = NEW(u32, u->offsets[ne] * 2);
  for (u32 i = 0; i < ne; ++i)
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 t = u->adj[j], d = u->dirs[j];
      g->adj[j * 2 + 0] = et[t * 3 + ((d + 1) % 3)];
      g->adj[j * 2 + 1] = et[t * 3 + ((d + 2) % 3)];
    }
  return g;
}

/* get_edge_tet_star (star.cpp:202-278): triangle star, then the edge
 * opposite to this one in every adjacent tet */
static ograph* star_edge_tet(orc_mesh* m)
{
  ograph* tg = star_edge_tri(m);
  u32 ne = mesh_count(m, 1);
  u32 const* er = ask_down(m, 3, 1);
  oup const* u = ask_up(m, 1, 3);
  u32* deg = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i)
    deg[i] = (tg->offsets[i + 1] - tg->offsets[i]) + (u->offsets[i + 1] - u->offsets[i]);
  ograph* g = NEW(ograph, 1);
  g->offsets = u32_exscan(deg, ne);
  free(deg);
  g->adj = NEW(u32, g->offsets[ne]);
  u32 const* oe = opp(3, 1);
  for (u32 i = 0; i < ne; ++i) {
    u32 o = g->offsets[i];
    for (u32 j = tg->offsets[i]; j < tg->offsets[i + 1]; ++j)
      g->adj[o++] = tg->adj[j];
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j)
      g->adj[o++] = er[u->adj[j] * 6 + oe[u->dirs[j]]];
  }
  free(tg->offsets);
  free(tg->adj);
  free(tg);
  return g;
}

/* mesh_get_star (star.cpp:280-315) */
static ograph* get_star(orc_mesh* m, u32 ld, u32 hd)
{
  if (ld == 0 && hd == m->dim && m->dim != 1 && mesh_has_dim(m, 1))
    return get_star(m, 0, 1);
  if (ld == 0 && hd == 1)
    return star_vertex_edge(m);
  if (ld == 1 && hd == 2)
    return star_edge_tri(m);
  if (ld == 1 && hd == 3 && mesh_has_dim(m, 2))
    return star_edge_tet(m);
  return star_general(m, ld, hd);
}

/* element_dual_from_verts (dual.cpp:40-98): neighbour across each side
 * found by intersecting the element lists of the side's vertices */
static u32* dual_from_verts(orc_mesh* m)
{
  u32 dim = m->dim, sd = dim - 1;
  u32 ne = mesh_count(m, dim);
  u32 const* ev = ask_down(m, dim, 0);
  oup const* u = ask_up(m, 0, dim);
  u32 vpe = dim + 1, spe = down_deg[dim][sd], vps = down_deg[sd][0];
  u32* out = NEW(u32, ne * spe);
  u32 buf[MAXUP];
  for (u32 i = 0; i < ne; ++i)
    for (u32 j = 0; j < spe; ++j) {
      u32 const* sc = CANON(dim, sd, 0, j);
      u32 n = 0;
      for (u32 k = 0; k < vps; ++k) {
        u32 v = ev[i * vpe + sc[k]];
        u32 f = u->offsets[v], e = u->offsets[v + 1];
        if (k) {
          u32 nn = 0;
          for (u32 a = 0; a < n; ++a) {
            int has = 0;
            for (u32 b = f; b < e; ++b)
              if (u->adj[b] == buf[a])
                has = 1;
            if (has)
              buf[nn++] = buf[a];
          }
          n = nn;
        } else {
          if (e - f > MAXUP)
            die("dual_from_verts: vertex degree above MAX_UP");
          for (u32 b = f; b < e; ++b)
            if (u->adj[b] != i)
              buf[n++] = u->adj[b];
        }
      }
      if (n > 1)
        die("dual_from_verts: non-manifold side");
      out[i * spe + j] = n ? buf[0] : INVALID;
    }
  return out;
}

/* element_dual_from_sides (dual.cpp:116-156) */
static u32* dual_from_sides(orc_mesh* m)
{
  u32 dim = m->dim;
  u32 ne = mesh_count(m, dim);
  u32 spe = down_deg[dim][dim - 1];
  u32 const* se = ask_down(m, dim, dim - 1);
  oup const* u = ask_up(m, dim - 1, dim);
  u32* out = NEW(u32, ne * spe);
  for (u32 i = 0; i < ne; ++i)
    for (u32 j = 0; j < spe; ++j) {
      u32 s = se[i * spe + j];
      u32 k, o = INVALID;
      for (k = u->offsets[s]; k < u->offsets[s + 1]; ++k)
        if (u->adj[k] != i) {
          o = u->adj[k];
          break;
        }
      out[i * spe + j] = o;
    }
  return out;
}

/* bridge_graph_general (bridge_graph.cpp:12-80): one edge per graph
 * neighbour pair (i, adj) with i < adj, numbered by i then adjacency
 * order; direction = position in i's adjacency list */
static void bridge_graph(u32 n, u32 const* offs, u32 const* adj, u32* p_nedges,
    u32** p_ev, u32** p_dirs)
{
  u32* deg = NEW(u32, n);
  for (u32 i = 0; i < n; ++i) {
    deg[i] = 0;
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      if (i < adj[j])
        deg[i]++;
  }
  u32* bo = u32_exscan(deg, n);
  free(deg);
  u32 ne = bo[n];
  u32* ev = NEW(u32, ne * 2);
  u32* dirs = p_dirs ? NEW(u32, ne) : 0;
  for (u32 i = 0; i < n; ++i) {
    u32 e = bo[i];
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      if (i < adj[j]) {
        ev[e * 2 + 0] = i;
        ev[e * 2 + 1] = adj[j];
        if (dirs)
          dirs[e] = j - offs[i];
        ++e;
      }
  }
  free(bo);
  *p_nedges = ne;
  *p_ev = ev;
  if (p_dirs)
    *p_dirs = dirs;
}

/* mesh_ask_down (mesh.cpp:136-181) */
static u32 const* ask_down(orc_mesh* m, u32 h, u32 l)
{
  if (h > m->dim || l > h)
    die("ask_down: bad dimensions");
  if (m->down[h][l])
    return m->down[h][l];
  if (l == h) {
    u32 n = m->counts[h];
    u32* id = NEW(u32, n);
    for (u32 i = 0; i < n; ++i)
      id[i] = i;
    m->down[h][l] = id;
  } else if (l > 0) {
    m->down[h][l] = reflect_down(m, h, l);
  } else {
    if (m->full)
      die("ask_down: missing entities in a full mesh");
    if (h == 1 && m->dim == 3) {
      /* 3D edges from the vertex-to-vertex star across tets */
      ograph const* g = ask_star(m, 0, m->dim);
      u32 ne;
      u32* ev;
      bridge_graph(m->counts[0], g->offsets, g->adj, &ne, &ev, 0);
      mesh_set_ents(m, 1, ne, ev);
    } else {
      /* sides: bridge_dual_graph + derive_sides (bridge_graph.cpp:82-108,
       * derive_sides.cpp:10-52) */
      if (h != m->dim - 1)
        die("ask_down: cannot derive this dimension");
      u32 const* dual = ask_dual(m);
      u32 dim = m->dim, ne = m->counts[dim];
      u32 spe = down_deg[dim][dim - 1];
      u32* offs = NEW(u32, ne + 1);
      for (u32 i = 0; i <= ne; ++i)
        offs[i] = i * spe;
      u32 ns;
      u32* es;
      u32* ess;
      bridge_graph(ne, offs, dual, &ns, &es, &ess);
      free(offs);
      u32 const* ev = m->down[dim][0];
      u32 vps = down_deg[dim - 1][0], vpe = dim + 1;
      u32* sv = NEW(u32, ns * vps);
      for (u32 i = 0; i < ns; ++i) {
        u32 e = es[i * 2];
        u32 const* sc = CANON(dim, dim - 1, 0, ess[i]);
        for (u32 j = 0; j < vps; ++j)
          sv[i * vps + j] = ev[e * vpe + sc[j]];
      }
      free(es);
      free(ess);
      mesh_set_ents(m, h, ns, sv);
    }
  }
  return m->down[h][l];
}

/* mesh_ask_up (mesh.cpp:190-217) */
static oup const* ask_up(orc_mesh* m, u32 l, u32 h)
{
  if (m->up[l][h])
    return m->up[l][h];
  if (l == h) {
    u32 n = m->counts[h];
    oup* u = NEW(oup, 1);
    u->offsets = NEW(u32, n + 1);
    for (u32 i = 0; i <= n; ++i)
      u->offsets[i] = i;
    u->adj = u32_copy(u->offsets, n);
    u->dirs = u32_filled(n, 0);
    m->up[l][h] = u;
  } else {
    u32 const* lh = ask_down(m, h, l);
    m->up[l][h] = up_from_down(h, l, m->counts[h], m->counts[l], lh);
  }
  return m->up[l][h];
}

/* mesh_ask_star (mesh.cpp:226-244) */
static ograph const* ask_star(orc_mesh* m, u32 l, u32 h)
{
  if (m->star[l][h])
    return m->star[l][h];
  if (l == h) {
    ograph* g = NEW(ograph, 1);
    g->offsets = u32_filled(m->counts[l] + 1, 0);
    g->adj = NEW(u32, 1);
    m->star[l][h] = g;
  } else {
    m->star[l][h] = get_star(m, l, h);
  }
  return m->star[l][h];
}

/* mesh_ask_dual (mesh.cpp:252-261) */
static u32 const* ask_dual(orc_mesh* m)
{
  if (!m->dual) {
    if (mesh_has_dim(m, m->dim - 1))
      m->dual = dual_from_sides(m);
    else
      m->dual = dual_from_verts(m);
  }
  return m->dual;
}

/* ------------------------------------------------------------- marks
 * mark.cpp:13-266 */
static u32* mark_down(orc_mesh* m, u32 hd, u32 ld, u32 const* marked_highs)
{
  u32 nl = mesh_count(m, ld);
  oup const* u = ask_up(m, ld, hd);
  u32* out = NEW(u32, nl);
  for (u32 i = 0; i < nl; ++i) {
    out[i] = 0;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j)
      if (marked_highs[u->adj[j]]) {
        out[i] = 1;
        break;
      }
  }
  return out;
}

static u32* mark_up(orc_mesh* m, u32 ld, u32 hd, u32 const* marked_lows)
{
  u32 nh = mesh_count(m, hd);
  u32 lph = down_deg[hd][ld];
  u32 const* lh = ask_down(m, hd, ld);
  u32* out = NEW(u32, nh);
  for (u32 i = 0; i < nh; ++i) {
    out[i] = 0;
    for (u32 j = 0; j < lph; ++j)
      if (marked_lows[lh[i * lph + j]]) {
        out[i] = 1;
        break;
      }
  }
  return out;
}

/* mesh_mark_dual_layers (mark.cpp:90-126) */
static void mark_dual_layers(orc_mesh* m, u32** marked, u32 nlayers)
{
  u32 dim = m->dim;
  u32 ne = mesh_count(m, dim);
  u32 const* dual = ask_dual(m);
  u32 deg = down_deg[dim][dim - 1];
  for (u32 l = 0; l < nlayers; ++l) {
    u32* in = *marked;
    u32* out = NEW(u32, ne);
    for (u32 i = 0; i < ne; ++i) {
      out[i] = in[i];
      for (u32 j = 0; j < deg; ++j) {
        u32 o = dual[i * deg + j];
        if (o != INVALID && in[o])
          out[i] = 1;
      }
    }
    free(in);
    *marked = out;
  }
}

/* mesh_mark_class (mark.cpp:128-169) */
static u32* mark_class(orc_mesh* m, u32 ent_dim, u32 target_dim, u32 target_id)
{
  u32 n = mesh_count(m, ent_dim);
  u32 const* cd = (u32 const*) find_tag(m, ent_dim, "class_dim")->data;
  u32 const* ci = 0;
  if (target_id != INVALID)
    ci = (u32 const*) find_tag(m, ent_dim, "class_id")->data;
  u32* out = NEW(u32, n);
  for (u32 i = 0; i < n; ++i) {
    out[i] = 0;
    if (cd[i] == target_dim)
      if (target_id == INVALID || ci[i] == target_id)
        out[i] = 1;
  }
  return out;
}

/* mesh_unmark_boundary (mark.cpp:180-200) */
static void unmark_boundary(orc_mesh* m, u32 ent_dim, u32* marked)
{
  u32 const* cd = (u32 const*) find_tag(m, ent_dim, "class_dim")->data;
  u32 n = mesh_count(m, ent_dim);
  for (u32 i = 0; i < n; ++i)
    if (cd[i] != m->dim)
      marked[i] = 0;
}

/* mesh_mark_slivers (mark.cpp:202-232) */
static u32* mark_slivers(orc_mesh* m, double good_qual, u32 nlayers)
{
  u32 ne = mesh_count(m, m->dim);
  double* q = mesh_qualities(m);
  u32* s = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i)
    s[i] = (q[i] < good_qual) ? 1 : 0;
  free(q);
  mark_dual_layers(m, &s, nlayers);
  return s;
}

/* mesh_mark_part_boundary (mark.cpp:234-258) */
static u32* mark_part_boundary(orc_mesh* m)
{
  u32 dim = m->dim;
  u32 ns = mesh_count(m, dim - 1);
  oup const* u = ask_up(m, dim - 1, dim);
  u32* out = NEW(u32, ns);
  for (u32 i = 0; i < ns; ++i)
    out[i] = (u->offsets[i + 1] - u->offsets[i]) < 2;
  return out;
}

/* ------------------------------------------------ model classification
 * derive_model.cpp:24-448, bfs.cpp:18-103 (setup only) */
static void derive_class_dim(orc_mesh* m, double crease_angle)
{
  u32 dim = m->dim;
  u32 ne = mesh_count(m, dim);
  tag_add(&m->tags[dim], T_U32, "class_dim", 1, u32_filled(ne, dim));
  u32* bs = mark_part_boundary(m);
  u32 ns = mesh_count(m, dim - 1);
  u32* scd = NEW(u32, ns);
  for (u32 i = 0; i < ns; ++i)
    scd[i] = dim - bs[i];
  tag_add(&m->tags[dim - 1], T_U32, "class_dim", 1, scd);
  /* boundary side normals (derive_model.cpp:24-81) */
  double const* x = coords_of(m);
  u32 const* sv = ask_down(m, dim - 1, 0);
  double* nrm = NEW(double, 3 * ns);
  for (u32 i = 0; i < ns; ++i) {
    if (!bs[i])
      continue;
    double* n = nrm + i * 3;
    if (dim - 1 == 1) {
      sub3(x + 3 * sv[i * 2 + 1], x + 3 * sv[i * 2 + 0], n);
    } else {
      double b0[3], b1[3];
      sub3(x + 3 * sv[i * 3 + 1], x + 3 * sv[i * 3 + 0], b0);
      sub3(x + 3 * sv[i * 3 + 2], x + 3 * sv[i * 3 + 0], b1);
      cross(b0, b1, n);
    }
    double s = 1.0 / sqrt(dot3(n, n)); /* normalize_vector, algebra.hpp:133-139 */
    for (u32 k = 0; k < 3; ++k)
      n[k] = n[k] * s;
  }
  u32* bh = mark_down(m, dim - 1, dim - 2, bs);
  u32 nh = mesh_count(m, dim - 2);
  oup const* hs = ask_up(m, dim - 2, dim - 1);
  u32* crease = NEW(u32, nh);
  for (u32 i = 0; i < nh; ++i) {
    double angle = 0;
    if (bh[i]) {
      u32 nadj = 0, adj[2] = {0, 0};
      for (u32 j = hs->offsets[i]; j < hs->offsets[i + 1]; ++j) {
        u32 s = hs->adj[j];
        if (bs[s]) {
          if (nadj >= 2)
            die("derive_model: more than two boundary sides at a hinge");
          adj[nadj++] = s;
        }
      }
      angle = acos(dot3(nrm + adj[0] * 3, nrm + adj[1] * 3));
    }
    crease[i] = angle > crease_angle;
  }
  free(bs);
  free(nrm);
  u32* hcd = NEW(u32, nh);
  for (u32 i = 0; i < nh; ++i)
    hcd[i] = dim - bh[i] - crease[i];
  free(bh);
  tag_add(&m->tags[dim - 2], T_U32, "class_dim", 1, hcd);
  if (dim == 2) {
    free(crease);
    return;
  }
  u32 nv = mesh_count(m, 0);
  oup const* ve = ask_up(m, 0, 1);
  u32* vcd = NEW(u32, nv);
  for (u32 i = 0; i < nv; ++i) {
    u32 nin = 0, d = 3;
    for (u32 j = ve->offsets[i]; j < ve->offsets[i + 1]; ++j) {
      u32 e = ve->adj[j];
      if (crease[e])
        ++nin;
      if (hcd[e] < d)
        d = hcd[e];
    }
    vcd[i] = d - (nin > 2);
  }
  free(crease);
  tag_add(&m->tags[0], T_U32, "class_dim", 1, vcd);
}

/* connected_components via bfs_full (bfs.cpp:18-103): components are
 * numbered in order of their lowest node */
static void connected_components(u32 n, u32 const* offs, u32 const* adj, u32* comp)
{
  u32* layer = u32_filled(n, INVALID);
  u32* queue = NEW(u32, n);
  u32 begin = 0, end = 0, c = 0;
  for (u32 i = 0; i < n; ++i) {
    if (layer[i] != INVALID)
      continue;
    queue[end++] = i;
    layer[i] = 0;
    comp[i] = c;
    while (begin != end) {
      u32 u = queue[begin++];
      for (u32 j = offs[u]; j < offs[u + 1]; ++j) {
        u32 v = adj[j];
        if (layer[v] == INVALID) {
          queue[end++] = v;
          layer[v] = layer[u] + 1;
          comp[v] = comp[u];
        }
      }
    }
    ++c;
  }
  free(layer);
  free(queue);
}

/* set_equal_order_class_id (derive_model.cpp:270-396) */
static void equal_order_class_id(orc_mesh* m, u32 d)
{
  u32 n = mesh_count(m, d);
  u32* eq = mark_class(m, d, d, INVALID);
  u32* eqo = u32_exscan(eq, n);
  free(eq);
  u32 neq = eqo[n];
  u32* br = mark_class(m, d - 1, d, INVALID);
  u32 bpe = down_deg[d][d - 1];
  u32 const* be = ask_down(m, d, d - 1);
  oup const* u = ask_up(m, d - 1, d);
  u32* deg = u32_filled(neq, 0);
  for (u32 i = 0; i < n; ++i) {
    if (eqo[i] == eqo[i + 1])
      continue;
    u32 k = 0;
    for (u32 j = 0; j < bpe; ++j)
      if (br[be[i * bpe + j]])
        ++k;
    deg[eqo[i]] = k;
  }
  u32* offs = u32_exscan(deg, neq);
  free(deg);
  u32* adj = NEW(u32, offs[neq]);
  for (u32 i = 0; i < n; ++i) {
    if (eqo[i] == eqo[i + 1])
      continue;
    u32 k = offs[eqo[i]];
    for (u32 j = 0; j < bpe; ++j) {
      u32 b = be[i * bpe + j];
      if (!br[b])
        continue;
      u32 l;
      for (l = u->offsets[b]; l < u->offsets[b + 1]; ++l) {
        u32 o = u->adj[l];
        if (o == i)
          continue;
        if (eqo[o] != eqo[o + 1]) {
          adj[k++] = eqo[o];
          break;
        }
      }
      if (l == u->offsets[b + 1])
        die("derive_model: dangling bridge");
    }
  }
  free(br);
  u32* comp = NEW(u32, neq);
  connected_components(neq, offs, adj, comp);
  free(offs);
  free(adj);
  u32* cid = u32_filled(n, 0);
  for (u32 i = 0; i < n; ++i)
    if (eqo[i] != eqo[i + 1])
      cid[i] = comp[eqo[i]];
  free(comp);
  free(eqo);
  tag_add(&m->tags[d], T_U32, "class_id", 1, cid);
}

/* project_class_id_onto (derive_model.cpp:398-416) */
static void project_class_id(orc_mesh* m, u32 ld)
{
  u32 nl = mesh_count(m, ld);
  oup const* u = ask_up(m, ld, ld + 1);
  u32 const* hcd = (u32 const*) find_tag(m, ld + 1, "class_dim")->data;
  u32 const* hci = (u32 const*) find_tag(m, ld + 1, "class_id")->data;
  u32 const* lcd = (u32 const*) find_tag(m, ld, "class_dim")->data;
  u32* lci = (u32*) find_tag(m, ld, "class_id")->data;
  for (u32 i = 0; i < nl; ++i)
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 h = u->adj[j];
      if (hcd[h] == lcd[i])
        lci[i] = hci[h];
    }
}

/* mesh_derive_model (derive_model.cpp:418-448) */
void orc_derive_model(orc_mesh* m, double crease_angle)
{
  derive_class_dim(m, crease_angle);
  u32 nv = mesh_count(m, 0);
  u32 const* vcd = (u32 const*) find_tag(m, 0, "class_dim")->data;
  u32* vci = u32_filled(nv, 0);
  u32 nmv = 0;
  for (u32 i = 0; i < nv; ++i)
    if (vcd[i] == 0)
      vci[i] = nmv++;
  tag_add(&m->tags[0], T_U32, "class_id", 1, vci);
  for (u32 d = 1; d <= m->dim; ++d)
    equal_order_class_id(m, d);
  for (u32 dd = 1; dd <= m->dim; ++dd)
    project_class_id(m, m->dim - dd);
}

void orc_derive_model_full(orc_mesh* m)
{
  orc_derive_model(m, REF_PI / 4);
  m->full = 1;
}

void orc_set_full(orc_mesh* m) { m->full = 1; }

/* ---------------------------------------------------- independent set
 * indset.cpp:33-139: Jacobi rounds of a deterministic Luby variant;
 * goodness ties broken by the lower global number (= index, serial) */
enum { NOT_IN_SET = 0, IN_SET = 1, UNKNOWN = 2 };

static u32* find_indset(orc_mesh* m, u32 ent_dim, u32 const* filter,
    double const* goodness)
{
  u32 n = mesh_count(m, ent_dim);
  ograph const* g = ask_star(m, ent_dim, m->dim);
  u32* state = NEW(u32, n);
  for (u32 i = 0; i < n; ++i)
    state[i] = filter[i] ? UNKNOWN : NOT_IN_SET;
  u32* old = NEW(u32, n);
  for (u32 it = 0; it < indset_max_rounds; ++it) {
    memcpy(old, state, sizeof(u32) * n);
    for (u32 i = 0; i < n; ++i) {
      if (old[i] != UNKNOWN)
        continue;
      u32 f = g->offsets[i], e = g->offsets[i + 1];
      int excluded = 0;
      for (u32 j = f; j < e; ++j)
        if (old[g->adj[j]] == IN_SET) {
          state[i] = NOT_IN_SET;
          excluded = 1;
          break;
        }
      if (excluded)
        continue;
      double myg = goodness[i];
      int wait = 0;
      for (u32 j = f; j < e; ++j) {
        u32 o = g->adj[j];
        if (old[o] == NOT_IN_SET)
          continue;
        double og = goodness[o];
        if (myg == og && o < i) {
          wait = 1;
          break;
        }
        if (myg < og) {
          wait = 1;
          break;
        }
      }
      if (!wait)
        state[i] = IN_SET;
    }
    if (u32_max(state, n) < UNKNOWN) {
      free(old);
      return state;
    }
  }
  die("independent set did not converge in the round limit");
  return 0;
}

void orc_find_indset(orc_mesh* m, unsigned ent_dim, unsigned const* candidates,
    double const* goodness, unsigned* out)
{
  u32 n = mesh_count(m, ent_dim);
  u32* s = find_indset(m, ent_dim, candidates, goodness);
  memcpy(out, s, sizeof(u32) * n);
  free(s);
}

/* ---------------------------------------------------------- inherit
 * inherit.cpp:14-232: class tags of products come from their domains */
static u32 prods_per_dom(u32 dom_dim, u32 src_dim, u32 prod_dim) /* refine_topology.cpp:110-121 */
{
  if (prod_dim == 0)
    return (dom_dim == src_dim) ? 1 : 0;
  u32 base_dim = prod_dim - 1;
  u32 opp_dim = dom_dim - 1 - base_dim;
  if (src_dim < opp_dim)
    return 0;
  return down_deg[src_dim][opp_dim];
}

typedef struct {
  u32 ndoms[4];
  u32* offs[4]; /* prods_of_doms_offsets; owned where noted */
  int own[4];
} prods_t;

static void prods_free(prods_t* p)
{
  for (u32 i = 0; i < 4; ++i)
    if (p->own[i])
      free(p->offs[i]);
}

static void ngen_offsets(prods_t const* p, u32 go[5]) /* inherit.cpp:14-36 */
{
  go[0] = 0;
  for (u32 i = 0; i < 4; ++i) {
    u32 ng = (p->ndoms[i] && p->offs[i]) ? p->offs[i][p->ndoms[i]] : 0;
    go[i + 1] = go[i] + ng;
  }
}

/* setup_refine (inherit.cpp:104-138), MESH_FULL branch and reduced branch */
static void setup_refine(orc_mesh* m, u32 src_dim, u32 prod_dim,
    u32* const dom_offsets[4], prods_t* p)
{
  memset(p, 0, sizeof(*p));
  u32 np = mesh_count(m, prod_dim);
  p->ndoms[0] = np;
  p->offs[0] = u32_negate_offsets(dom_offsets[prod_dim], np);
  p->own[0] = 1;
  u32 lo, hi;
  if (m->full) {
    lo = (src_dim > prod_dim) ? src_dim : prod_dim;
    hi = m->dim;
  } else {
    lo = hi = (prod_dim == 0) ? src_dim : m->dim;
  }
  for (u32 d = lo; d <= hi; ++d) {
    p->ndoms[d] = mesh_count(m, d);
    p->offs[d] = u32_scale(dom_offsets[d], p->ndoms[d] + 1,
        prods_per_dom(d, src_dim, prod_dim));
    p->own[d] = 1;
  }
}

/* inherit_uint_tag / inherit_class (inherit.cpp:168-232) */
static void inherit_tag(orc_mesh* m, orc_mesh* mo, u32 prod_dim,
    prods_t const* p, char const* name)
{
  otag* t = find_tag(m, prod_dim, name);
  u32 const* dd[4];
  dd[0] = (u32 const*) t->data;
  for (u32 d = 1; d < 4; ++d) {
    otag* td = (d <= m->dim) ? find_tag(m, d, name) : 0;
    dd[d] = td ? (u32 const*) td->data : 0;
  }
  u32 go[5];
  ngen_offsets(p, go);
  u32 w = t->ncomps;
  u32* out = NEW(u32, go[4] * w);
  for (u32 i = 0; i < 4; ++i) {
    if (!(p->ndoms[i] && dd[i] && p->offs[i]))
      continue;
    u32* g = (u32*) expand(dd[i], sizeof(u32), p->offs[i], p->ndoms[i], w);
    memcpy(out + go[i] * w, g, sizeof(u32) * (go[i + 1] - go[i]) * w);
    free(g);
  }
  tag_add(&mo->tags[prod_dim], T_U32, name, w, out);
}

static void inherit_class(orc_mesh* m, orc_mesh* mo, u32 prod_dim, prods_t const* p)
{
  if (!find_tag(m, 0, "class_dim"))
    return;
  if (!m->full)
    die("inherit_class needs a full mesh");
  inherit_tag(m, mo, prod_dim, p, "class_dim");
  if (find_tag(m, 0, "class_id"))
    inherit_tag(m, mo, prod_dim, p, "class_id");
}

/* ------------------------------------------------- solution transfer
 * Element fields tagged OSH_TRANSFER_CONSERVE or OSH_TRANSFER_POINTWISE
 * follow every operation; every other element tag is dropped by it. */
#define TR_POINTWISE 1
#define TR_CONSERVE 2
#define MAX_NCOMPS 32 /* coarsen_conserve.cpp:11 */
#define MAX_PTS 16    /* qr.hpp:6 */

/* element_sizes / measure_entity (size.cpp:14-55, size.hpp:48-57) */
static double* element_sizes(orc_mesh* m)
{
  u32 d = m->dim, n = mesh_count(m, d), vpe = d + 1;
  u32 const* ev = ask_down(m, d, 0);
  double const* x = coords_of(m);
  double* out = NEW(double, n);
  for (u32 i = 0; i < n; ++i) {
    double ex[4][3];
    for (u32 j = 0; j < vpe; ++j)
      memcpy(ex[j], x + 3 * ev[i * vpe + j], sizeof(double) * 3);
    if (d == 3)
      out[i] = tet_volume((double const (*)[3]) ex);
    else if (d == 2)
      out[i] = tri_area((double const (*)[3]) ex);
    else
      out[i] = sqrt(vsqdist(ex[0], ex[1]));
  }
  return out;
}

/* interp_to_elems / average_element_field (element_field.cpp:10-45,
 * element_field.hpp:8-25): the mean of the vertex values */
static double* interp_to_elems(orc_mesh* m, double const* in, u32 nc)
{
  u32 d = m->dim, n = mesh_count(m, d), vpe = d + 1;
  u32 const* ev = ask_down(m, d, 0);
  double* out = NEW(double, (size_t) n * nc);
  for (u32 i = 0; i < n; ++i) {
    double* o = out + (size_t) i * nc;
    for (u32 j = 0; j < nc; ++j)
      o[j] = 0;
    for (u32 k = 0; k < vpe; ++k)
      for (u32 j = 0; j < nc; ++j)
        o[j] += in[(size_t) ev[i * vpe + k] * nc + j];
    for (u32 j = 0; j < nc; ++j)
      o[j] /= vpe;
  }
  return out;
}

static int has_elem_tags(orc_mesh* m, int transfer)
{
  otags* ts = &m->tags[m->dim];
  for (u32 i = 0; i < ts->n; ++i)
    if (ts->at[i]->type == T_F64 && (transfer < 0 || ts->at[i]->transfer == transfer))
      return 1;
  return 0;
}

/* the element tag of the output: the same elements' values (offs_same over
 * n old elements), then ngen generated values */
static double* same_then_gen(otag const* t, u32 const* soff, u32 n, u32 ngen)
{
  u32 w = t->ncomps, nsame = soff[n];
  double* out = NEW(double, (size_t)(nsame + ngen) * w);
  double const* in = (double const*) t->data;
  for (u32 i = 0; i < n; ++i)
    if (soff[i] != soff[i + 1])
      memcpy(out + (size_t) soff[i] * w, in + (size_t) i * w, sizeof(double) * w);
  return out;
}

/* refine_conserve / refine_fit (refine_conserve.cpp:10-71,
 * refine_fit.cpp:10-70), in that order (refine_common.cpp:110-113): the
 * same elements keep their values; the children of a split element share
 * its value equally (conserve) or copy it (pointwise) */
static void refine_transfer(orc_mesh* m, orc_mesh* mo, prods_t const* p)
{
  u32 dim = m->dim;
  u32 go[5];
  ngen_offsets(p, go);
  static int const kinds[2] = {TR_CONSERVE, TR_POINTWISE};
  for (u32 kk = 0; kk < 2; ++kk)
    for (u32 ti = 0; ti < m->tags[dim].n; ++ti) {
      otag* t = m->tags[dim].at[ti];
      if (t->type != T_F64 || t->transfer != kinds[kk])
        continue;
      u32 w = t->ncomps;
      double const* in = (double const*) t->data;
      double* out = NEW(double, (size_t) go[4] * w);
      double* same = (double*) expand(in, sizeof(double), p->offs[0], p->ndoms[0], w);
      memcpy(out + (size_t) go[0] * w, same, sizeof(double) * (go[1] - go[0]) * w);
      free(same);
      u32 const* po = p->offs[dim];
      for (u32 i = 0; p->ndoms[dim] && po && i < p->ndoms[dim]; ++i) {
        u32 f = po[i], e = po[i + 1];
        double denom = e - f;
        for (u32 j = f; j < e; ++j)
          for (u32 k = 0; k < w; ++k) {
            double v = in[(size_t) i * w + k];
            out[(size_t)(go[dim] + j) * w + k] = (kinds[kk] == TR_CONSERVE) ? v / denom : v;
          }
      }
      tag_add2(&mo->tags[dim], T_F64, t->name, w, kinds[kk], out);
    }
}

static double sign_of(double x) { return (x < 0) ? -1 : 1; } /* qr.cpp:10-13 */

/* get_reflector2 (qr.cpp:40-62) */
static u32 get_reflector2(double a[MAX_PTS][4], double v[MAX_PTS], u32 k, u32 npts)
{
  double cnorm = 0;
  for (u32 i = k; i < npts; ++i)
    cnorm += a[i][k] * a[i][k];
  cnorm = sqrt(cnorm);
  if (cnorm < 1e-6)
    return 0;
  for (u32 i = 0; i < k; ++i)
    v[i] = 0;
  for (u32 i = k; i < npts; ++i)
    v[i] = a[i][k];
  v[k] += sign_of(a[k][k]) * cnorm;
  double rnorm = 0;
  for (u32 i = k; i < npts; ++i)
    rnorm += v[i] * v[i];
  rnorm = sqrt(rnorm);
  for (u32 i = k; i < npts; ++i)
    v[i] /= rnorm;
  return 1;
}

/* qr_decomp2 (qr.cpp:151-168) with reflect_columns2 / reflect_rows2
 * (qr.cpp:76-86, 100-110): Householder QR of the npts x 4 matrix a; Q is
 * accumulated explicitly; returns the number of reflections taken */
static u32 qr_decomp2(double a[MAX_PTS][4], double q[MAX_PTS][MAX_PTS], double r[MAX_PTS][4],
    u32 npts)
{
  for (u32 i = 0; i < npts; ++i)
    for (u32 j = 0; j < 4; ++j)
      r[i][j] = a[i][j];
  for (u32 i = 0; i < npts; ++i)
    for (u32 j = 0; j < npts; ++j)
      q[i][j] = (double) (i == j);
  double v[MAX_PTS] = {0};
  u32 rank = 0;
  for (u32 k = 0; k < 4; ++k)
    if (get_reflector2(r, v, k, npts)) {
      for (u32 j = 0; j < 4; ++j) {
        double dot = 0;
        for (u32 i = k; i < npts; ++i)
          dot += r[i][j] * v[i];
        for (u32 i = k; i < npts; ++i)
          r[i][j] -= 2 * dot * v[i];
      }
      for (u32 i = 0; i < npts; ++i) {
        double dot = 0;
        for (u32 j = k; j < npts; ++j)
          dot += q[i][j] * v[j];
        for (u32 j = k; j < npts; ++j)
          q[i][j] -= 2 * dot * v[j];
      }
      ++rank;
    }
  return rank;
}

/* qr_solve2 (qr.cpp:170-200): y = Q^T b over the first 4 columns, then
 * back substitution with R */
static void qr_solve2(double q[MAX_PTS][MAX_PTS], double r[MAX_PTS][4], double const* b,
    u32 npts, double x[4])
{
  double y[4];
  for (u32 i = 0; i < 4; ++i) {
    y[i] = 0;
    for (u32 j = 0; j < npts; ++j)
      y[i] += q[j][i] * b[j];
  }
  for (u32 ii = 0; ii < 4; ++ii) {
    u32 i = 4 - ii - 1;
    x[i] = y[i];
    for (u32 j = i + 1; j < 4; ++j)
      x[i] -= r[i][j] * x[j];
    x[i] /= r[i][i];
  }
}

/* coarsen_conserve (coarsen_conserve.cpp:15-139) and coarsen_fit
 * (coarsen_fit.cpp:17-187) for the elements of a coarsen: cavity = the
 * old elements around a collapsing vertex.  Conserve: the cavity's total
 * is split over its new elements by new size.  Pointwise (3D only, the
 * reference asserts): a least-squares linear fit over the first MAX_PTS
 * old element centroids, evaluated at each new centroid; fewer than 4
 * points or a rank-deficient fit give the cavity average.  The fit leaves
 * an element "coordinates" tag (the centroids) on the old mesh, computed
 * only if absent, and on the new mesh, before the pointwise tags. */
static void coarsen_transfer(orc_mesh* m, orc_mesh* mo, u32 const* gen_off_v, u32 const* goff,
    u32 const* soff)
{
  u32 dim = m->dim, nv = mesh_count(m, 0), n = mesh_count(m, dim);
  u32 nsame = soff[n], ngen = goff[n];
  oup const* up = ask_up(m, 0, dim);
  if (has_elem_tags(m, -1)) {
    double* sizes = element_sizes(mo);
    for (u32 ti = 0; ti < m->tags[dim].n; ++ti) {
      otag* t = m->tags[dim].at[ti];
      if (t->type != T_F64 || t->transfer != TR_CONSERVE)
        continue;
      u32 w = t->ncomps;
      if (w > MAX_NCOMPS)
        die("coarsen_conserve: too many components");
      double const* in = (double const*) t->data;
      double* out = same_then_gen(t, soff, n, ngen);
      double* gen = out + (size_t) nsame * w;
      for (u32 v = 0; v < nv; ++v) {
        if (gen_off_v[v] == gen_off_v[v + 1])
          continue;
        u32 f = up->offsets[v], e = up->offsets[v + 1];
        double sum[MAX_NCOMPS] = {0};
        for (u32 j = f; j < e; ++j)
          for (u32 k = 0; k < w; ++k)
            sum[k] = sum[k] + in[(size_t) up->adj[j] * w + k];
        double cav = 0;
        for (u32 j = f; j < e; ++j) {
          u32 el = up->adj[j];
          if (goff[el] != goff[el + 1])
            cav += sizes[goff[el] + nsame];
        }
        for (u32 j = f; j < e; ++j) {
          u32 el = up->adj[j];
          if (goff[el] == goff[el + 1])
            continue;
          double s = sizes[goff[el] + nsame] / cav;
          for (u32 k = 0; k < w; ++k)
            gen[(size_t) goff[el] * w + k] = sum[k] * s;
        }
      }
      tag_add2(&mo->tags[dim], T_F64, t->name, w, TR_CONSERVE, out);
    }
    free(sizes);
  }
  if (!has_elem_tags(m, TR_POINTWISE))
    return;
  if (dim != 3)
    die("coarsen_fit: pointwise element fields need a 3D mesh");
  if (!find_tag(m, dim, "coordinates"))
    tag_add(&m->tags[dim], T_F64, "coordinates", 3, interp_to_elems(m, coords_of(m), 3));
  tag_add(&mo->tags[dim], T_F64, "coordinates", 3, interp_to_elems(mo, coords_of(mo), 3));
  double const* oc = (double const*) find_tag(m, dim, "coordinates")->data;
  double const* nc = (double const*) find_tag(mo, dim, "coordinates")->data;
  for (u32 ti = 0; ti < m->tags[dim].n; ++ti) {
    otag* t = m->tags[dim].at[ti];
    if (t->type != T_F64 || t->transfer != TR_POINTWISE)
      continue;
    u32 w = t->ncomps;
    if (w > MAX_NCOMPS)
      die("coarsen_fit: too many components");
    double const* in = (double const*) t->data;
    double* out = same_then_gen(t, soff, n, ngen);
    double* gen = out + (size_t) nsame * w;
    for (u32 v = 0; v < nv; ++v) {
      if (gen_off_v[v] == gen_off_v[v + 1])
        continue;
      u32 f = up->offsets[v], e = up->offsets[v + 1];
      u32 npts = e - f;
      if (npts > MAX_PTS)
        npts = MAX_PTS;
      int fit = npts >= 4;
      double q[MAX_PTS][MAX_PTS], r[MAX_PTS][4];
      if (fit) {
        double a[MAX_PTS][4];
        for (u32 j = 0; j < npts; ++j) {
          u32 el = up->adj[f + j];
          a[j][0] = 1.0;
          memcpy(&a[j][1], oc + 3 * (size_t) el, sizeof(double) * 3);
        }
        if (qr_decomp2(a, q, r, npts) < 4)
          fit = 0;
      }
      for (u32 k = 0; k < w; ++k) {
        double c[4], avg = 0;
        if (fit) {
          double b[MAX_PTS];
          for (u32 j = 0; j < npts; ++j)
            b[j] = in[(size_t) up->adj[f + j] * w + k];
          qr_solve2(q, r, b, npts, c);
        } else {
          for (u32 j = f; j < e; ++j)
            avg += in[(size_t) up->adj[j] * w + k];
          avg /= (e - f);
        }
        for (u32 j = f; j < e; ++j) {
          u32 el = up->adj[j];
          if (goff[el] == goff[el + 1])
            continue;
          u32 ge = goff[el];
          if (fit) {
            double const* xc = nc + 3 * (size_t)(ge + nsame);
            double d = 0;
            for (u32 l = 0; l < 3; ++l)
              d += c[1 + l] * xc[l];
            gen[(size_t) ge * w + k] = c[0] + d;
          } else {
            gen[(size_t) ge * w + k] = avg;
          }
        }
      }
    }
    tag_add2(&mo->tags[dim], T_F64, t->name, w, TR_POINTWISE, out);
  }
}

/* swap_conserve (swap_conserve.cpp:15-131) and swap_fit (swap_fit.cpp:
 * 16-127): cavity = the old elements around a swapped edge; conserve
 * splits the cavity total by new size, pointwise takes the cavity average
 * (3D only, the reference asserts) */
static void swap_transfer(orc_mesh* m, orc_mesh* mo, u32 const* goff_e, u32 const* soff)
{
  u32 dim = m->dim, ne = mesh_count(m, 1), n = mesh_count(m, dim);
  u32 nsame = soff[n], ngen = goff_e[ne];
  oup const* up = ask_up(m, 1, dim);
  if (has_elem_tags(m, -1)) {
    double* sizes = element_sizes(mo);
    for (u32 ti = 0; ti < m->tags[dim].n; ++ti) {
      otag* t = m->tags[dim].at[ti];
      if (t->type != T_F64 || t->transfer != TR_CONSERVE)
        continue;
      u32 w = t->ncomps;
      if (w > MAX_NCOMPS)
        die("swap_conserve: too many components");
      double const* in = (double const*) t->data;
      double* out = same_then_gen(t, soff, n, ngen);
      double* gen = out + (size_t) nsame * w;
      for (u32 i = 0; i < ne; ++i) {
        if (goff_e[i] == goff_e[i + 1])
          continue;
        double sum[MAX_NCOMPS] = {0};
        for (u32 j = up->offsets[i]; j < up->offsets[i + 1]; ++j)
          for (u32 k = 0; k < w; ++k)
            sum[k] = sum[k] + in[(size_t) up->adj[j] * w + k];
        double cav = 0;
        for (u32 j = goff_e[i]; j < goff_e[i + 1]; ++j)
          cav += sizes[nsame + j];
        for (u32 j = goff_e[i]; j < goff_e[i + 1]; ++j) {
          double s = sizes[nsame + j] / cav;
          for (u32 k = 0; k < w; ++k)
            gen[(size_t) j * w + k] = sum[k] * s;
        }
      }
      tag_add2(&mo->tags[dim], T_F64, t->name, w, TR_CONSERVE, out);
    }
    free(sizes);
  }
  if (!has_elem_tags(m, TR_POINTWISE))
    return;
  if (dim != 3)
    die("swap_fit: pointwise element fields need a 3D mesh");
  for (u32 ti = 0; ti < m->tags[dim].n; ++ti) {
    otag* t = m->tags[dim].at[ti];
    if (t->type != T_F64 || t->transfer != TR_POINTWISE)
      continue;
    u32 w = t->ncomps;
    double const* in = (double const*) t->data;
    double* out = same_then_gen(t, soff, n, ngen);
    double* gen = out + (size_t) nsame * w;
    for (u32 i = 0; i < ne; ++i) {
      if (goff_e[i] == goff_e[i + 1])
        continue;
      u32 f = up->offsets[i], e = up->offsets[i + 1];
      for (u32 k = 0; k < w; ++k) {
        double avg = 0;
        for (u32 j = f; j < e; ++j)
          avg += in[(size_t) up->adj[j] * w + k];
        avg /= (e - f);
        for (u32 g = goff_e[i]; g < goff_e[i + 1]; ++g)
          gen[(size_t) g * w + k] = avg;
      }
    }
    tag_add2(&mo->tags[dim], T_F64, t->name, w, TR_POINTWISE, out);
  }
}

/* ------------------------------------------------------------ refine */

/* refine_quality (refine_qualities.cpp:16-82): worst child quality of
 * every candidate split; rejected splits clear their candidate flag */
static double* refine_qualities(orc_mesh* m, u32 src_dim, u32* cands,
    double qual_floor, int require_better)
{
  u32 dim = m->dim;
  u32 ns = mesh_count(m, src_dim);
  u32 const* vs = ask_down(m, src_dim, 0);
  u32 const* ve = ask_down(m, dim, 0);
  oup const* u = ask_up(m, src_dim, dim);
  double const* x = coords_of(m);
  double* eq = require_better ? mesh_qualities(m) : 0;
  double const* mt = metric_of(m);
  u32 base_dim = dim - 1;
  u32 vps = down_deg[src_dim][0], vpe = dim + 1;
  u32 const* bo = opp(dim, 0);
  double* out = NEW(double, ns);
  for (u32 i = 0; i < ns; ++i) {
    out[i] = 0;
    if (!cands[i])
      continue;
    double split[3] = {0, 0, 0}; /* average_element_field, element_field.hpp:8-25 */
    for (u32 k = 0; k < vps; ++k)
      for (u32 c = 0; c < 3; ++c)
        split[c] += x[vs[i * vps + k] * 3 + c];
    for (u32 c = 0; c < 3; ++c)
      split[c] /= vps;
    double msplit[6] = {0, 0, 0, 0, 0, 0};
    if (mt) {
      for (u32 k = 0; k < vps; ++k)
        for (u32 c = 0; c < 6; ++c)
          msplit[c] += mt[vs[i * vps + k] * 6 + c];
      for (u32 c = 0; c < 6; ++c)
        msplit[c] /= vps;
    }
    double minq = 1, old_minq = 1;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 dir = u->dirs[j], e = u->adj[j];
      if (eq && eq[e] < old_minq)
        old_minq = eq[e];
      u32 const* evs = CANON(dim, src_dim, 0, dir);
      for (u32 k = 0; k < vps; ++k) {
        u32 base = bo[evs[k]];
        u32 const* bv = CANON(dim, base_dim, 0, base);
        double ex[4][3], em[4][6];
        for (u32 l = 0; l + 1 < vpe; ++l) {
          memcpy(ex[l], x + 3 * ve[e * vpe + bv[l]], sizeof(double) * 3);
          if (mt)
            memcpy(em[l], mt + 6 * ve[e * vpe + bv[l]], sizeof(double) * 6);
        }
        memcpy(ex[vpe - 1], split, sizeof(double) * 3);
        memcpy(em[vpe - 1], msplit, sizeof(double) * 6);
        double q = corner_quality(dim, (double const (*)[3]) ex, (double const (*)[6]) em, mt, 0);
        if (!(q > 0))
          die("refine_quality: non-positive child quality");
        if (q < minq)
          minq = q;
      }
    }
    if ((minq < qual_floor) || (require_better && (minq <= old_minq)))
      cands[i] = 0;
    else
      out[i] = minq;
  }
  free(eq);
  return out;
}

/* choose_refinement_indset (refine_common.cpp:131-160) */
static int choose_refinement_indset(orc_mesh* m, u32 src_dim, double qual_floor,
    int require_better)
{
  u32 ns = mesh_count(m, src_dim);
  otag* ct = find_tag(m, src_dim, "candidate");
  if (!u32_max((u32 const*) ct->data, ns)) {
    tag_remove(&m->tags[src_dim], "candidate");
    return 0;
  }
  u32* good = u32_copy((u32 const*) ct->data, ns);
  tag_remove(&m->tags[src_dim], "candidate");
  double* q = refine_qualities(m, src_dim, good, qual_floor, require_better);
  if (!u32_max(good, ns)) {
    free(good);
    free(q);
    return 0;
  }
  u32* indset = find_indset(m, src_dim, good, q);
  free(good);
  free(q);
  tag_add(&m->tags[src_dim], T_U32, "indset", 1, indset);
  return 1;
}

/* refine_verts (refine_common.cpp:29-54) + refine_node
 * (refine_nodal.cpp:10-46): every F64 vertex tag gets the average of the
 * split entity's vertices appended, in split order */
static void refine_verts(orc_mesh* m, orc_mesh* mo, u32 src_dim, u32 const* gen_off)
{
  u32 nv = mesh_count(m, 0);
  u32 ns = mesh_count(m, src_dim);
  u32 nsplit = gen_off[ns];
  mesh_set_ents(mo, 0, nv + nsplit, 0);
  u32 const* vs = ask_down(m, src_dim, 0);
  u32 vps = down_deg[src_dim][0];
  for (u32 t = 0; t < m->tags[0].n; ++t) {
    otag* tg = m->tags[0].at[t];
    if (tg->type != T_F64)
      continue;
    u32 w = tg->ncomps;
    double const* f = (double const*) tg->data;
    double* out = NEW(double, (size_t)(nv + nsplit) * w);
    memcpy(out, f, sizeof(double) * nv * w);
    for (u32 i = 0; i < ns; ++i) {
      if (gen_off[i] == gen_off[i + 1])
        continue;
      double* o = out + (size_t)(nv + gen_off[i]) * w;
      for (u32 c = 0; c < w; ++c)
        o[c] = 0;
      for (u32 k = 0; k < vps; ++k)
        for (u32 c = 0; c < w; ++c)
          o[c] += f[vs[i * vps + k] * w + c];
      for (u32 c = 0; c < w; ++c)
        o[c] /= vps;
    }
    tag_add(&mo->tags[0], T_F64, tg->name, w, out);
  }
}

/* split_to_domain (splits_to_domains.cpp:13-56): a domain splits along
 * its first split source entity */
static void splits_to_domains(orc_mesh* m, u32 dom_dim, u32 src_dim,
    u32 const* off_src, u32 const* vert_src, u32** p_off, u32** p_dir, u32** p_vert)
{
  u32 nd = mesh_count(m, dom_dim);
  u32 const* sd = ask_down(m, dom_dim, src_dim);
  u32 spd = down_deg[dom_dim][src_dim];
  u32* will = u32_filled(nd, 0);
  u32* dir = u32_filled(nd, 0);
  u32* vert = u32_filled(nd, 0);
  for (u32 i = 0; i < nd; ++i)
    for (u32 j = 0; j < spd; ++j) {
      u32 s = sd[i * spd + j];
      if (off_src[s] == off_src[s + 1])
        continue;
      will[i] = 1;
      dir[i] = j;
      vert[i] = vert_src[s];
      break;
    }
  *p_off = u32_exscan(will, nd);
  free(will);
  *p_dir = dir;
  *p_vert = vert;
}

/* refine_domain_entity (refine_topology.cpp:26-107): products = the base
 * opposite each of the source's opposite-dim entities, plus the new vertex */
static void refine_topology(orc_mesh* m, u32 dom_dim, u32 src_dim, u32 prod_dim,
    u32 const* off, u32 const* dir, u32 const* vert, u32* out)
{
  u32 base_dim = prod_dim - 1;
  u32 opp_dim = dom_dim - 1 - base_dim;
  if (src_dim < opp_dim)
    return;
  u32 nd = mesh_count(m, dom_dim);
  if (!off[nd])
    return;
  u32 ops = down_deg[src_dim][opp_dim];
  u32 vpp = down_deg[prod_dim][0], vpb = vpp - 1, vpd = down_deg[dom_dim][0];
  u32 const* vd = ask_down(m, dom_dim, 0);
  u32 const* bo = opp(dom_dim, opp_dim);
  for (u32 i = 0; i < nd; ++i) {
    if (off[i] == off[i + 1])
      continue;
    u32 const* os = CANON(dom_dim, src_dim, opp_dim, dir[i]);
    u32* p = out + (size_t) off[i] * ops * vpp;
    for (u32 j = 0; j < ops; ++j) {
      u32 base = bo[os[j]];
      u32 const* bv = CANON(dom_dim, base_dim, 0, base);
      for (u32 k = 0; k < vpb; ++k)
        p[k] = vd[i * vpd + bv[k]];
      p[vpb] = vert[i];
      p += vpp;
    }
  }
}

/* refine_ents (refine_common.cpp:64-129) */
static void refine_ents(orc_mesh* m, orc_mesh* mo, u32 src_dim, u32 const* gen_off)
{
  u32 dim = m->dim;
  u32 ns = mesh_count(m, src_dim);
  u32 nv = mesh_count(m, 0);
  u32* gvs = u32_filled(ns, 0);
  for (u32 i = 0; i < ns; ++i)
    if (gen_off[i] != gen_off[i + 1])
      gvs[i] = nv + gen_off[i];
  u32* off[4] = {0, 0, 0, 0};
  u32* dir[4] = {0, 0, 0, 0};
  u32* vert[4] = {0, 0, 0, 0};
  off[0] = u32_filled(nv + 1, 0);
  for (u32 d = 1; d <= dim; ++d) {
    if (!mesh_has_dim(m, d))
      continue;
    if (d >= src_dim)
      splits_to_domains(m, d, src_dim, gen_off, gvs, &off[d], &dir[d], &vert[d]);
    else
      off[d] = u32_filled(mesh_count(m, d) + 1, 0);
  }
  free(gvs);
  for (u32 pd = 0; pd <= dim; ++pd) {
    if (!m->full && 0 < pd && pd < dim)
      continue;
    prods_t p;
    setup_refine(m, src_dim, pd, off, &p);
    u32 go[5];
    ngen_offsets(&p, go);
    if (pd) {
      u32 np = go[4];
      u32 vpp = down_deg[pd][0];
      u32* vo = NEW(u32, (size_t) np * vpp);
      u32 const* vp = ask_down(m, pd, 0);
      for (u32 i = 0; i < p.ndoms[0]; ++i)
        for (u32 j = p.offs[0][i]; j < p.offs[0][i + 1]; ++j)
          memcpy(vo + (size_t) j * vpp, vp + (size_t) i * vpp, sizeof(u32) * vpp);
      for (u32 d = 1; d <= dim; ++d)
        if (p.ndoms[d] && p.offs[d])
          refine_topology(m, d, src_dim, pd, off[d], dir[d], vert[d],
              vo + (size_t) go[d] * vpp);
      mesh_set_ents(mo, pd, np, vo);
    }
    inherit_class(m, mo, pd, &p);
    if (pd == dim)
      refine_transfer(m, mo, &p);
    prods_free(&p);
  }
  for (u32 d = 0; d <= dim; ++d) {
    free(off[d]);
    free(dir[d]);
    free(vert[d]);
  }
}

static void say(char const* verb, u32 n, u32 d)
{
  static char const* const names[4][2] = {{"vertex", "vertices"},
    {"edge", "edges"}, {"triangle", "triangles"}, {"tet", "tets"}};
  if (g_verbose)
    printf("%s %10u %s\n", verb, n, names[d][n != 1]);
}

/* refine_common (refine_common.cpp:162-190) */
static u32 refine_common(orc_mesh* m, u32 src_dim, double qual_floor, int require_better)
{
  if (!choose_refinement_indset(m, src_dim, qual_floor, require_better))
    return 0;
  otag* it = find_tag(m, src_dim, "indset");
  u32 ns = mesh_count(m, src_dim);
  u32 total = u32_sum((u32 const*) it->data, ns);
  u32* gen_off = u32_exscan((u32 const*) it->data, ns);
  tag_remove(&m->tags[src_dim], "indset");
  orc_mesh* mo = mesh_new(m->dim, m->full);
  refine_verts(m, mo, src_dim, gen_off);
  refine_ents(m, mo, src_dim, gen_off);
  free(gen_off);
  say("split", total, src_dim);
  mesh_overwrite(m, mo);
  return 1;
}

/* refine_by_size (refine.cpp:11-27) */
unsigned orc_refine_by_size(orc_mesh* m, double qual_floor)
{
  double* s = measure_edges(m);
  u32 ne = mesh_count(m, 1);
  u32* c = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i)
    c[i] = s[i] > 1.0;
  free(s);
  tag_add(&m->tags[1], T_U32, "candidate", 1, c);
  return refine_common(m, 1, qual_floor, 0);
}

/* uniformly_refine (refine.cpp:29-35) */
void orc_uniformly_refine(orc_mesh* m)
{
  u32 ne = mesh_count(m, 1);
  tag_add(&m->tags[1], T_U32, "candidate", 1, u32_filled(ne, 1));
  refine_common(m, 1, 0.0, 0);
}

/* ----------------------------------------------------------- coarsen */
enum { DONT_COLLAPSE = 0, COLLAPSE_BOTH = 3 }; /* collapse_codes.hpp:8-31 */
static u32 collapses(u32 code, u32 dir) { return (code & (1u << dir)) != 0; }
static u32 do_collapse(u32 code, u32 dir) { return code | (1u << dir); }
static u32 dont_collapse(u32 code, u32 dir) { return code & ~(1u << dir); }

/* check_basic_edge_class (check_collapse_class.cpp:12-55) */
static void check_basic_class(orc_mesh* m, u32* codes)
{
  u32 ne = mesh_count(m, 1);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vcd = (u32 const*) find_tag(m, 0, "class_dim")->data;
  u32 const* ecd = (u32 const*) find_tag(m, 1, "class_dim")->data;
  for (u32 i = 0; i < ne; ++i) {
    if (codes[i] == DONT_COLLAPSE)
      continue;
    u32 cd[3] = {vcd[ev[i * 2]], vcd[ev[i * 2 + 1]], ecd[i]};
    if (cd[0] == cd[1]) {
      if (cd[0] != cd[2])
        codes[i] = DONT_COLLAPSE;
      continue;
    }
    for (u32 j = 0; j < 2; ++j) {
      if (!collapses(codes[i], j))
        continue;
      if (cd[j] != cd[2])
        codes[i] = dont_collapse(codes[i], j);
    }
  }
}

/* check_for_exposed_side (check_collapse_class.cpp:57-140) */
static void check_exposed_sides(orc_mesh* m, u32 ent_dim, u32* codes)
{
  oup const* u = ask_up(m, 1, ent_dim);
  u32 const* vent = ask_down(m, ent_dim, 0);
  u32 vpe = down_deg[ent_dim][0];
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* sent = ask_down(m, ent_dim, ent_dim - 1);
  u32 spe = down_deg[ent_dim][ent_dim - 1];
  u32 const* ecd = (u32 const*) find_tag(m, ent_dim, "class_dim")->data;
  u32 const* scd = (u32 const*) find_tag(m, ent_dim - 1, "class_dim")->data;
  u32 const* so = opp(ent_dim, 0);
  u32 ne = mesh_count(m, 1);
  for (u32 i = 0; i < ne; ++i) {
    if (codes[i] == DONT_COLLAPSE)
      continue;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 ent = u->adj[j], dir = u->dirs[j];
      u32 const* eve = CANON(ent_dim, 1, 0, dir);
      for (u32 k = 0; k < 2; ++k) {
        if (!collapses(codes[i], k))
          continue;
        u32 va = ev[i * 2 + (1 - k)];
        u32 l;
        for (l = 0; l < 2; ++l) {
          u32 vb = vent[ent * vpe + eve[l]];
          if (va == vb) {
            u32 side = sent[ent * spe + so[eve[l]]];
            if (scd[side] != ecd[ent])
              codes[i] = dont_collapse(codes[i], k);
            break;
          }
        }
        if (l == 2)
          die("check_for_exposed_side: edge vertex not in entity");
      }
    }
  }
}

/* coarsen_quality (coarsen_qualities.cpp:12-79): per collapse direction,
 * worst quality of the elements around the collapsing vertex that do not
 * contain the target vertex, with the collapsing vertex moved onto it */
static double* coarsen_qualities(orc_mesh* m, u32* codes, double qual_floor,
    int require_better)
{
  u32 dim = m->dim;
  u32 ne = mesh_count(m, 1);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vel = ask_down(m, dim, 0);
  oup const* u = ask_up(m, 0, dim);
  double const* x = coords_of(m);
  double const* mt = metric_of(m);
  double* eq = require_better ? element_qualities(dim, mesh_count(m, dim), vel, x, mt) : 0;
  u32 vpe = dim + 1;
  u32 const* bo = opp(dim, 0);
  double* out = NEW(double, ne * 2);
  for (u32 i = 0; i < 2 * ne; ++i)
    out[i] = 0;
  for (u32 i = 0; i < ne; ++i) {
    if (codes[i] == DONT_COLLAPSE)
      continue;
    for (u32 j = 0; j < 2; ++j) {
      if (!collapses(codes[i], j))
        continue;
      double minq = 1, old_minq = 1;
      u32 col = ev[i * 2 + j], gen = ev[i * 2 + (1 - j)];
      double ex[4][3], em[4][6];
      memcpy(ex[vpe - 1], x + 3 * gen, sizeof(double) * 3);
      if (mt)
        memcpy(em[vpe - 1], mt + 6 * gen, sizeof(double) * 6);
      for (u32 k = u->offsets[col]; k < u->offsets[col + 1]; ++k) {
        u32 e = u->adj[k];
        u32 const* ve = vel + e * vpe;
        if (eq && eq[e] < old_minq)
          old_minq = eq[e];
        int adj = 0;
        for (u32 l = 0; l < vpe; ++l)
          if (ve[l] == gen)
            adj = 1;
        if (adj)
          continue;
        u32 const* bv = CANON(dim, dim - 1, 0, bo[u->dirs[k]]);
        for (u32 l = 0; l + 1 < vpe; ++l) {
          memcpy(ex[l], x + 3 * ve[bv[l]], sizeof(double) * 3);
          if (mt)
            memcpy(em[l], mt + 6 * ve[bv[l]], sizeof(double) * 6);
        }
        double q = corner_quality(dim, (double const (*)[3]) ex, (double const (*)[6]) em, mt, 1);
        if (q < minq)
          minq = q;
      }
      if ((minq < qual_floor) || (require_better && (minq <= old_minq)))
        codes[i] = dont_collapse(codes[i], j);
      else
        out[i * 2 + j] = minq;
    }
  }
  free(eq);
  return out;
}

/* collapse_to_vert (collapses_to_verts.cpp:13-70) */
static void collapses_to_verts(orc_mesh* m, u32 const* codes, double const* cq,
    u32** p_cand, double** p_vq)
{
  u32 nv = mesh_count(m, 0);
  oup const* u = ask_up(m, 0, 1);
  u32* cand = NEW(u32, nv);
  double* vq = NEW(double, nv);
  for (u32 i = 0; i < nv; ++i) {
    double maxq = 0;
    u32 is = 0;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 e = u->adj[j], d = u->dirs[j];
      if (!collapses(codes[e], d))
        continue;
      double q = cq[e * 2 + d];
      if (!is) {
        is = 1;
        maxq = q;
      } else if (q > maxq) {
        maxq = q;
      }
    }
    cand[i] = is;
    vq[i] = is ? maxq : 0;
  }
  *p_cand = cand;
  *p_vq = vq;
}

/* vertex_dest (collapses_to_verts.cpp:72-135): the first collapsing edge
 * (in up(0,1) order) whose quality equals the vertex maximum, with the
 * 1e10 -> -1e10 substitution of the reference */
static u32* vertex_destinations(orc_mesh* m, u32 const* codes, double const* cq,
    u32 const* indset, double const* vq)
{
  u32 nv = mesh_count(m, 0);
  u32 const* ev = ask_down(m, 1, 0);
  oup const* u = ask_up(m, 0, 1);
  u32* out = u32_filled(nv, INVALID);
  for (u32 i = 0; i < nv; ++i) {
    if (!indset[i])
      continue;
    double maxq = vq[i];
    if (maxq == 1e10)
      maxq = -1e10;
    u32 g = INVALID;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j) {
      u32 e = u->adj[j], d = u->dirs[j];
      if (!collapses(codes[e], d))
        continue;
      double q = cq[e * 2 + d];
      if (q == 1e10)
        q = -1e10;
      if (q == maxq) {
        g = ev[e * 2 + (1 - d)];
        break;
      }
    }
    if (g == INVALID)
      die("vertex_dest: no destination");
    out[i] = g;
  }
  return out;
}

/* collapse_to_ent (collapses_to_ents.cpp:11-115) */
static void collapses_to_ents(orc_mesh* m, u32 ed, u32 const* gen_off_v,
    u32 const* gen_vert_v, u32 const* fused_ents, u32** p_gen_off, u32** p_gen_vert,
    u32** p_gen_dir, u32** p_same_off, u32** p_fused_sides)
{
  u32 n = mesh_count(m, ed);
  u32 const* ve = ask_down(m, ed, 0);
  u32 const* se = ask_down(m, ed, ed - 1);
  u32 ns = mesh_count(m, ed - 1);
  u32* fused = u32_filled(ns, 0);
  u32 vpe = down_deg[ed][0], spe = down_deg[ed][ed - 1];
  u32 const* so = opp(ed, 0);
  u32* will = u32_filled(n, 0);
  u32* gv = u32_filled(n, 0);
  u32* gd = u32_filled(n, 0);
  u32* same = NEW(u32, n);
  for (u32 i = 0; i < n; ++i) {
    u32 const* v = ve + i * vpe;
    u32 col = INVALID, dir = INVALID;
    for (u32 j = 0; j < vpe; ++j)
      if (gen_off_v[v[j]] != gen_off_v[v[j] + 1]) {
        col = v[j];
        dir = j;
      }
    if (dir == INVALID) {
      same[i] = 1;
      continue;
    }
    same[i] = 0;
    u32 g = gen_vert_v[col];
    u32 gdir = INVALID;
    for (u32 j = 0; j < vpe; ++j)
      if (v[j] == g)
        gdir = j;
    if (gdir != INVALID) {
      fused[se[i * spe + so[gdir]]] = 1;
      continue;
    }
    if (!(fused_ents && fused_ents[i])) {
      will[i] = 1;
      gv[i] = g;
      gd[i] = dir;
    }
  }
  *p_gen_off = u32_exscan(will, n);
  free(will);
  *p_gen_vert = gv;
  *p_gen_dir = gd;
  *p_same_off = u32_exscan(same, n);
  free(same);
  *p_fused_sides = fused;
}

/* coarsen_ents (coarsen_common.cpp:35-89) + coarsen_elem
 * (coarsen_topology.cpp:9-62) + concat_verts_of_ents (inherit.cpp:234-255) */
static void coarsen_ents(orc_mesh* m, orc_mesh* mo, u32 ed, u32 const* gen_off_v,
    u32 const* gen_vert_v, u32 const* same_off_v, u32 const* fused_ents,
    u32** p_fused_sides)
{
  u32 n = mesh_count(m, ed);
  u32 const* ve = ask_down(m, ed, 0);
  u32 *goff, *gvert, *gdir, *soff;
  collapses_to_ents(m, ed, gen_off_v, gen_vert_v, fused_ents, &goff, &gvert, &gdir,
      &soff, p_fused_sides);
  u32 vpe = down_deg[ed][0];
  u32 ngen = goff[n], nsame = soff[n];
  u32 const* bo = opp(ed, 0);
  u32* vout = NEW(u32, (size_t)(nsame + ngen) * vpe);
  for (u32 i = 0; i < n; ++i) {
    if (soff[i] != soff[i + 1])
      memcpy(vout + (size_t) soff[i] * vpe, ve + (size_t) i * vpe, sizeof(u32) * vpe);
    if (goff[i] == goff[i + 1])
      continue;
    u32* g = vout + (size_t)(nsame + goff[i]) * vpe;
    u32 const* bv = CANON(ed, ed - 1, 0, bo[gdir[i]]);
    for (u32 j = 0; j + 1 < vpe; ++j)
      g[j] = ve[i * vpe + bv[j]];
    g[vpe - 1] = gvert[i];
  }
  free(gvert);
  free(gdir);
  /* remap_conn (coarsen_common.cpp:29-33) */
  for (size_t k = 0; k < (size_t)(nsame + ngen) * vpe; ++k)
    vout[k] = same_off_v[vout[k]];
  mesh_set_ents(mo, ed, nsame + ngen, vout);
  /* setup_coarsen (inherit.cpp:140-156): survivors, then products of dim ed */
  prods_t p;
  memset(&p, 0, sizeof(p));
  p.ndoms[0] = n;
  p.offs[0] = soff;
  p.ndoms[ed] = n;
  p.offs[ed] = goff;
  inherit_class(m, mo, ed, &p);
  if (ed == m->dim)
    coarsen_transfer(m, mo, gen_off_v, goff, soff);
  free(goff);
  free(soff);
}

/* tags_subset (subset.cpp:15-43) for vertices */
static void tags_subset_verts(orc_mesh* m, orc_mesh* mo, u32 const* offs)
{
  u32 nv = mesh_count(m, 0);
  for (u32 t = 0; t < m->tags[0].n; ++t) {
    otag* tg = m->tags[0].at[t];
    void* d = expand(tg->data, tag_esz(tg->type), offs, nv, tg->ncomps);
    tag_add(&mo->tags[0], tg->type, tg->name, tg->ncomps, d);
  }
}

/* coarsen_common (coarsen_common.cpp:109-239) */
static u32 coarsen_common(orc_mesh* m, double qual_floor, int require_better)
{
  u32 dim = m->dim;
  u32 ne = mesh_count(m, 1);
  otag* ct = find_tag(m, 1, "col_codes");
  if (u32_max((u32 const*) ct->data, ne) == DONT_COLLAPSE) {
    tag_remove(&m->tags[1], "col_codes");
    return 0;
  }
  u32* codes = u32_copy((u32 const*) ct->data, ne);
  tag_remove(&m->tags[1], "col_codes");
  check_basic_class(m, codes);
  for (u32 d = 2; d <= dim; ++d)
    check_exposed_sides(m, d, codes);
  if (u32_max(codes, ne) == DONT_COLLAPSE) {
    free(codes);
    return 0;
  }
  double* cq = coarsen_qualities(m, codes, qual_floor, require_better);
  if (u32_max(codes, ne) == DONT_COLLAPSE) {
    free(codes);
    free(cq);
    return 0;
  }
  u32* cand;
  double* vq;
  collapses_to_verts(m, codes, cq, &cand, &vq);
  u32* indset = find_indset(m, 0, cand, vq);
  free(cand);
  /* coarsen_interior (coarsen_common.cpp:191-219) */
  u32 nv = mesh_count(m, 0);
  u32* gvv = vertex_destinations(m, codes, cq, indset, vq);
  free(codes);
  free(cq);
  free(vq);
  u32 total = u32_sum(indset, nv);
  u32* goff = u32_exscan(indset, nv);
  free(indset);
  u32* soff = u32_negate_offsets(goff, nv);
  orc_mesh* mo = mesh_new(dim, m->full);
  mesh_set_ents(mo, 0, soff[nv], 0);
  tags_subset_verts(m, mo, soff);
  u32* fused[4] = {0, 0, 0, 0};
  for (u32 dd = 0; dd < dim; ++dd) {
    u32 d = dim - dd;
    coarsen_ents(m, mo, d, goff, gvv, soff, fused[d], &fused[d - 1]);
  }
  for (u32 d = 0; d <= dim; ++d)
    free(fused[d]);
  free(gvv);
  free(goff);
  free(soff);
  say("collapsed", total, 1);
  mesh_overwrite(m, mo);
  return 1;
}

/* coarsen_by_size (coarsen.cpp:13-36) */
unsigned orc_coarsen_by_size(orc_mesh* m, double qual_floor, double size_floor)
{
  u32 ne = mesh_count(m, 1);
  double* s = measure_edges(m);
  u32* c = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i)
    c[i] = (s[i] < size_floor) ? COLLAPSE_BOTH : DONT_COLLAPSE;
  free(s);
  tag_add(&m->tags[1], T_U32, "col_codes", 1, c);
  return coarsen_common(m, qual_floor, 0);
}

/* coarsen_slivers (coarsen.cpp:38-70) */
unsigned orc_coarsen_slivers(orc_mesh* m, double qual_floor, unsigned nlayers)
{
  u32 dim = m->dim;
  u32* sl = mark_slivers(m, qual_floor, nlayers);
  u32* mv = mark_down(m, dim, 0, sl);
  free(sl);
  u32 ne = mesh_count(m, 1);
  u32 const* ev = ask_down(m, 1, 0);
  u32* c = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i) {
    c[i] = 0;
    for (u32 j = 0; j < 2; ++j)
      if (mv[ev[i * 2 + j]])
        c[i] = do_collapse(c[i], j);
  }
  free(mv);
  tag_add(&m->tags[1], T_U32, "col_codes", 1, c);
  return coarsen_common(m, 0.0, 1);
}

/* -------------------------------------------------------------- swap
 * Triangulations of an N-gon (N = 3..7) and their unique triangles
 * (edge_swap.cpp:21-200).  Data that must match the reference exactly:
 * the triangulation index is the swap "code" and ties go to the lowest
 * code. */
#define MAX_EDGE_SWAP 7
static u32 const swap_mesh_sizes[8] = {0, 0, 0, 1, 2, 3, 4, 5};
static u32 const swap_mesh_counts[8] = {0, 0, 0, 1, 2, 5, 14, 42};
static u32 const sw_tris3[1][3] = {{0, 1, 2}};
static u32 const sw_mesh3[1] = {0};
static u32 const sw_tris4[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {1, 2, 3}};
static u32 const sw_mesh4[4] = {0, 1, 2, 3};
static u32 const sw_tris5[10][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 1, 4},
  {1, 3, 4}, {1, 2, 3}, {2, 3, 4}, {0, 2, 4}, {0, 1, 3}, {1, 2, 4}};
static u32 const sw_mesh5[15] = {0, 1, 2, 3, 4, 5, 0, 6, 7, 2, 5, 8, 3, 6, 9};
static u32 const sw_tris6[20][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5},
  {0, 2, 5}, {2, 4, 5}, {2, 3, 4}, {0, 3, 5}, {3, 4, 5}, {0, 2, 4}, {2, 3, 5},
  {1, 2, 3}, {0, 1, 3}, {0, 1, 5}, {1, 4, 5}, {1, 3, 4}, {0, 1, 4}, {1, 3, 5},
  {1, 2, 4}, {1, 2, 5}};
static u32 const sw_mesh6[56] = {0, 1, 2, 3, 0, 4, 5, 6, 0, 1, 7, 8, 0, 3, 6, 9,
  0, 4, 8, 10, 2, 3, 11, 12, 11, 13, 14, 15, 7, 8, 11, 12, 3, 11, 15, 16, 8, 11,
  13, 17, 6, 13, 14, 18, 3, 6, 16, 18, 5, 6, 13, 19, 8, 10, 13, 19};
static u32 const sw_tris7[35][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5},
  {0, 5, 6}, {0, 3, 6}, {3, 5, 6}, {3, 4, 5}, {0, 4, 6}, {4, 5, 6}, {0, 3, 5},
  {3, 4, 6}, {0, 2, 4}, {2, 3, 4}, {0, 2, 6}, {2, 5, 6}, {2, 4, 5}, {0, 2, 5},
  {2, 4, 6}, {2, 3, 5}, {2, 3, 6}, {0, 1, 3}, {1, 2, 3}, {0, 1, 4}, {1, 3, 4},
  {0, 1, 6}, {1, 5, 6}, {1, 4, 5}, {0, 1, 5}, {1, 4, 6}, {1, 3, 5}, {1, 3, 6},
  {1, 2, 4}, {1, 2, 5}, {1, 2, 6}};
static u32 const sw_mesh7[210] = {0, 1, 2, 3, 4, 0, 1, 5, 6, 7, 0, 1, 2, 8, 9,
  0, 1, 4, 7, 10, 0, 1, 5, 9, 11, 0, 3, 4, 12, 13, 0, 13, 14, 15, 16, 0, 8, 9,
  12, 13, 0, 4, 13, 16, 17, 0, 9, 13, 14, 18, 0, 7, 14, 15, 19, 0, 4, 7, 17, 19,
  0, 6, 7, 14, 20, 0, 9, 11, 14, 20, 2, 3, 4, 21, 22, 5, 6, 7, 21, 22, 2, 8, 9,
  21, 22, 4, 7, 10, 21, 22, 5, 9, 11, 21, 22, 3, 4, 22, 23, 24, 22, 24, 25, 26,
  27, 8, 9, 22, 23, 24, 4, 22, 24, 27, 28, 9, 22, 24, 25, 29, 7, 22, 25, 26, 30,
  4, 7, 22, 28, 30, 6, 7, 22, 25, 31, 9, 11, 22, 25, 31, 3, 4, 13, 23, 32, 13,
  25, 26, 27, 32, 8, 9, 13, 23, 32, 4, 13, 27, 28, 32, 9, 13, 25, 29, 32, 13, 16,
  25, 26, 33, 4, 13, 16, 28, 33, 13, 15, 16, 25, 34, 9, 13, 18, 25, 34, 7, 19,
  25, 26, 33, 4, 7, 19, 28, 33, 7, 15, 19, 25, 34, 6, 7, 20, 25, 34, 9, 11, 20,
  25, 34};
static u32 const (*const swap_tris[8])[3] = {0, 0, 0, sw_tris3, sw_tris4,
  sw_tris5, sw_tris6, sw_tris7};
static u32 const* const swap_meshes[8] = {0, 0, 0, sw_mesh3, sw_mesh4, sw_mesh5,
  sw_mesh6, sw_mesh7};

/* interior edges of every triangulation, derived the way the reference's
 * generator does (test_print_swap_edges.cpp:29-64): build the
 * triangulation as a reduced 2D mesh, derive its edges, keep the
 * non-boundary ones in edge order. */
static u32 swap_nint[8];
static u32* swap_int[8]; /* [code * nint * 2] */
static int swap_tables_ready = 0;

static void make_swap_tables(void)
{
  if (swap_tables_ready)
    return;
  for (u32 n = 4; n <= MAX_EDGE_SWAP; ++n) {
    u32 nm = swap_mesh_counts[n], nt = swap_mesh_sizes[n];
    for (u32 c = 0; c < nm; ++c) {
      orc_mesh* m = mesh_new(2, 0);
      mesh_set_ents(m, 0, n, 0);
      u32* conn = NEW(u32, nt * 3);
      for (u32 k = 0; k < nt; ++k)
        for (u32 l = 0; l < 3; ++l)
          conn[k * 3 + l] = swap_tris[n][swap_meshes[n][c * nt + k]][l];
      mesh_set_ents(m, 2, nt, conn);
      u32* bd = mark_part_boundary(m);
      u32 ne = mesh_count(m, 1);
      u32 nint = ne - u32_sum(bd, ne);
      if (c == 0) {
        swap_nint[n] = nint;
        swap_int[n] = NEW(u32, nm * nint * 2);
      } else if (nint != swap_nint[n]) {
        die("swap tables: inconsistent interior edge count");
      }
      u32 const* ev = ask_down(m, 1, 0);
      u32 l = 0;
      for (u32 k = 0; k < ne; ++k)
        if (!bd[k]) {
          swap_int[n][(c * nint + l) * 2 + 0] = ev[k * 2 + 0];
          swap_int[n][(c * nint + l) * 2 + 1] = ev[k * 2 + 1];
          ++l;
        }
      free(bd);
      orc_free(m);
    }
  }
  swap_tables_ready = 1;
}

unsigned orc_swap_nint_edges(unsigned ring_size)
{
  make_swap_tables();
  return swap_nint[ring_size];
}

void orc_swap_int_edges(unsigned ring_size, unsigned code, unsigned* out)
{
  make_swap_tables();
  u32 k = swap_nint[ring_size];
  memcpy(out, swap_int[ring_size] + code * k * 2, sizeof(u32) * k * 2);
}

/* choose_edge_swap (edge_swap.cpp:230-280): best triangulation of the
 * ring; each triangle makes two tets with the edge endpoints; worst tet
 * quality per triangle cached; strict '>' keeps the first best code */
static u32 choose_edge_swap(u32 rs, double const (*ex)[3], double const (*rx)[3],
    double* quality, double const (*em)[6], double const (*rm)[6])
{
  u32 tpm = swap_mesh_sizes[rs], nm = swap_mesh_counts[rs];
  u32 const* mesh = swap_meshes[rs];
  u32 const (*tris)[3] = swap_tris[rs];
  unsigned char cached[35];
  double cache[35];
  memset(cached, 0, sizeof(cached));
  u32 code = INVALID;
  double best = -1;
  for (u32 i = 0; i < nm; ++i) {
    double mesh_minq = 1;
    for (u32 j = 0; j < tpm; ++j) {
      u32 tri = mesh[j];
      if (!cached[tri]) {
        double tx[4][3], tm[4][6];
        for (u32 k = 0; k < 3; ++k) {
          memcpy(tx[k], rx[tris[tri][k]], sizeof(double) * 3);
          if (em)
            memcpy(tm[k], rm[tris[tri][k]], sizeof(double) * 6);
        }
        double tmin = 1;
        for (u32 k = 0; k < 2; ++k) {
          memcpy(tx[3], ex[1 - k], sizeof(double) * 3);
          if (em)
            memcpy(tm[3], em[1 - k], sizeof(double) * 6);
          double q = em ? metric_quality(3, (double const (*)[3]) tx, (double const (*)[6]) tm, 1)
                        : tet_quality((double const (*)[3]) tx);
          if (q < tmin)
            tmin = q;
          double t[3], tt[6];
          memcpy(t, tx[0], sizeof(t));
          memcpy(tx[0], tx[1], sizeof(t));
          memcpy(tx[1], t, sizeof(t));
          if (em) {
            memcpy(tt, tm[0], sizeof(tt));
            memcpy(tm[0], tm[1], sizeof(tt));
            memcpy(tm[1], tt, sizeof(tt));
          }
        }
        cached[tri] = 1;
        cache[tri] = tmin;
      }
      double tq = cache[tri];
      if (tq < mesh_minq)
        mesh_minq = tq;
      if (mesh_minq < 0)
        break;
    }
    if (mesh_minq > best) {
      code = i;
      best = mesh_minq;
    }
    mesh += tpm;
  }
  *quality = best;
  return code;
}

unsigned orc_choose_edge_swap(unsigned ring_size, double const* edge_x6,
    double const* ring_x, double* quality)
{
  return choose_edge_swap(ring_size, (double const (*)[3]) edge_x6,
      (double const (*)[3]) ring_x, quality, 0, 0);
}

/* find_edge_ring (edge_ring.cpp:12-74): the ring's vertices ordered by
 * chaining each tet's opposite edge (a -> b) by insertion */
static u32 find_edge_ring(orc_mesh* m, u32 edge, u32* edge_v, u32* ring_v)
{
  oup const* u = ask_up(m, 1, 3);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vt = ask_down(m, 3, 0);
  u32 f = u->offsets[edge], e = u->offsets[edge + 1];
  u32 rs = e - f;
  if (rs < 3)
    die("find_edge_ring: fewer than 3 tets around an edge");
  if (rs > MAX_EDGE_SWAP)
    return rs;
  u32 const* oe = opp(3, 1);
  edge_v[0] = ev[edge * 2 + 0];
  edge_v[1] = ev[edge * 2 + 1];
  u32 ra[MAX_EDGE_SWAP], rb[MAX_EDGE_SWAP];
  for (u32 i = f; i < e; ++i) {
    u32 t = u->adj[i], te = u->dirs[i];
    u32 const* tv = vt + t * 4;
    u32 const* tve = CANON(3, 1, 0, te);
    u32 align = (tv[tve[0]] != edge_v[0]) ? 1 : 0;
    u32 const* tvo = CANON(3, 1, 0, oe[te]);
    ra[i - f] = tv[tvo[align ^ 0]];
    rb[i - f] = tv[tvo[align ^ 1]];
  }
  for (u32 i = 0; i + 1 < rs; ++i) {
    u32 j;
    for (j = i + 1; j < rs; ++j)
      if (ra[j] == rb[i])
        break;
    if (j == rs)
      die("find_edge_ring: ring does not close");
    u32 ta = ra[j], tb = rb[j];
    ra[j] = ra[i + 1];
    rb[j] = rb[i + 1];
    ra[i + 1] = ta;
    rb[i + 1] = tb;
  }
  for (u32 i = 0; i < rs; ++i)
    ring_v[i] = ra[i];
  return rs;
}

/* swap_quality (swap_qualities.cpp:13-122) */
static void swap_qualities(orc_mesh* m, u32* cands, double** p_q, u32** p_rs)
{
  u32 ne = mesh_count(m, 1);
  oup const* u = ask_up(m, 1, 3);
  double const* x = coords_of(m);
  double const* mt = metric_of(m);
  double* eq = mesh_qualities(m);
  double* q = NEW(double, ne);
  u32* rsz = u32_filled(ne, 0);
  for (u32 i = 0; i < ne; ++i) {
    q[i] = 0;
    if (!cands[i])
      continue;
    double old_minq = 1;
    for (u32 j = u->offsets[i]; j < u->offsets[i + 1]; ++j)
      if (eq[u->adj[j]] < old_minq)
        old_minq = eq[u->adj[j]];
    u32 edge_v[2], ring_v[MAX_EDGE_SWAP];
    u32 rs = find_edge_ring(m, i, edge_v, ring_v);
    if (rs > MAX_EDGE_SWAP) {
      cands[i] = 0;
      continue;
    }
    double ex[2][3], rx[MAX_EDGE_SWAP][3];
    memcpy(ex[0], x + 3 * edge_v[0], sizeof(double) * 3);
    memcpy(ex[1], x + 3 * edge_v[1], sizeof(double) * 3);
    for (u32 j = 0; j < rs; ++j)
      memcpy(rx[j], x + 3 * ring_v[j], sizeof(double) * 3);
    double emt[2][6], rmt[MAX_EDGE_SWAP][6];
    if (mt) {
      memcpy(emt[0], mt + 6 * edge_v[0], sizeof(double) * 6);
      memcpy(emt[1], mt + 6 * edge_v[1], sizeof(double) * 6);
      for (u32 j = 0; j < rs; ++j)
        memcpy(rmt[j], mt + 6 * ring_v[j], sizeof(double) * 6);
    }
    double sq;
    choose_edge_swap(rs, (double const (*)[3]) ex, (double const (*)[3]) rx, &sq,
        mt ? (double const (*)[6]) emt : 0, mt ? (double const (*)[6]) rmt : 0);
    if (sq > old_minq) {
      q[i] = sq;
      rsz[i] = rs;
    } else {
      cands[i] = 0;
    }
  }
  free(eq);
  *p_q = q;
  *p_rs = rsz;
}

static u32 count_swap_ents(u32 rs, u32 d) /* edge_swap.cpp:375-386 */
{
  switch (d) {
    case 1: return swap_nint[rs];
    case 2: return swap_mesh_sizes[rs] + 2 * swap_nint[rs];
    case 3: return 2 * swap_mesh_sizes[rs];
  }
  die("count_swap_ents");
  return 0;
}

/* get_swap_tets / get_swap_edges / get_swap_tris (edge_swap.cpp:282-373) */
static void get_swap_ents(u32 rs, u32 code, u32 d, u32 const* edge_v,
    u32 const* ring_v, u32* out)
{
  u32 tpm = swap_mesh_sizes[rs];
  u32 const* mesh = swap_meshes[rs] + code * tpm;
  u32 const (*tris)[3] = swap_tris[rs];
  u32 nint = swap_nint[rs];
  u32 const* ie = nint ? swap_int[rs] + code * nint * 2 : 0;
  if (d == 3) {
    for (u32 i = 0; i < tpm; ++i) {
      u32 tv[4];
      for (u32 j = 0; j < 3; ++j)
        tv[j] = ring_v[tris[mesh[i]][j]];
      for (u32 j = 0; j < 2; ++j) {
        tv[3] = edge_v[1 - j];
        for (u32 k = 0; k < 4; ++k)
          *out++ = tv[k];
        u32 t = tv[0];
        tv[0] = tv[1];
        tv[1] = t;
      }
    }
  } else if (d == 1) {
    for (u32 i = 0; i < nint; ++i)
      for (u32 j = 0; j < 2; ++j)
        out[i * 2 + j] = ring_v[ie[i * 2 + j]];
  } else {
    for (u32 i = 0; i < nint; ++i)
      for (u32 j = 0; j < 2; ++j) {
        for (u32 k = 0; k < 2; ++k)
          out[i * 6 + j * 3 + k] = ring_v[ie[i * 2 + k]];
        out[i * 6 + j * 3 + 2] = edge_v[j];
      }
    out += nint * 6;
    for (u32 i = 0; i < tpm; ++i)
      for (u32 j = 0; j < 3; ++j)
        out[i * 3 + j] = ring_v[tris[mesh[i]][j]];
  }
}

/* swap_ents (swap.cpp:27-69) + swap_cavity (swap_topology.cpp:16-101) */
static void swap_ents(orc_mesh* m, orc_mesh* mo, u32 d, u32 const* indset,
    u32 const* rsz)
{
  u32 ne = mesh_count(m, 1);
  u32* cnt = NEW(u32, ne);
  for (u32 i = 0; i < ne; ++i)
    cnt[i] = indset[i] ? count_swap_ents(rsz[i], d) : 0;
  u32* goff = u32_exscan(cnt, ne);
  free(cnt);
  u32 ngen = goff[ne];
  u32 vpe = down_deg[d][0];
  u32* gen = NEW(u32, (size_t) ngen * vpe);
  double const* x = coords_of(m);
  double const* mt = metric_of(m);
  for (u32 i = 0; i < ne; ++i) {
    if (!indset[i])
      continue;
    u32 edge_v[2], ring_v[MAX_EDGE_SWAP + 1];
    u32 rs = find_edge_ring(m, i, edge_v, ring_v);
    if (rs > MAX_EDGE_SWAP)
      die("swap_cavity: ring too large");
    double ex[2][3], rx[MAX_EDGE_SWAP][3];
    memcpy(ex[0], x + 3 * edge_v[0], sizeof(double) * 3);
    memcpy(ex[1], x + 3 * edge_v[1], sizeof(double) * 3);
    for (u32 j = 0; j < rs; ++j)
      memcpy(rx[j], x + 3 * ring_v[j], sizeof(double) * 3);
    double emt[2][6], rmt[MAX_EDGE_SWAP][6];
    if (mt) {
      memcpy(emt[0], mt + 6 * edge_v[0], sizeof(double) * 6);
      memcpy(emt[1], mt + 6 * edge_v[1], sizeof(double) * 6);
      for (u32 j = 0; j < rs; ++j)
        memcpy(rmt[j], mt + 6 * ring_v[j], sizeof(double) * 6);
    }
    double sq;
    u32 code = choose_edge_swap(rs, (double const (*)[3]) ex, (double const (*)[3]) rx, &sq,
        mt ? (double const (*)[6]) emt : 0, mt ? (double const (*)[6]) rmt : 0);
    get_swap_ents(rs, code, d, edge_v, ring_v, gen + (size_t) goff[i] * vpe);
  }
  u32* old = mark_up(m, 1, d, indset);
  u32 n = mesh_count(m, d);
  for (u32 i = 0; i < n; ++i)
    old[i] = !old[i];
  u32* soff = u32_exscan(old, n);
  free(old);
  u32 const* ve = ask_down(m, d, 0);
  u32 nsame = soff[n];
  u32* same = (u32*) expand(ve, sizeof(u32), soff, n, vpe);
  u32* vout = (u32*) concat(same, gen, sizeof(u32), nsame, ngen, vpe);
  free(same);
  free(gen);
  mesh_set_ents(mo, d, nsame + ngen, vout);
  if (m->full) {
    /* setup_swap (inherit.cpp:158-166): products come from edges */
    prods_t p;
    memset(&p, 0, sizeof(p));
    p.ndoms[0] = n;
    p.offs[0] = soff;
    p.ndoms[1] = ne;
    p.offs[1] = goff;
    inherit_class(m, mo, d, &p);
  }
  if (d == m->dim)
    swap_transfer(m, mo, goff, soff);
  free(goff);
  free(soff);
}

/* swap_common + swap_interior (swap.cpp:71-136) */
static u32 swap_common(orc_mesh* m, u32* cands)
{
  u32 ne = mesh_count(m, 1);
  if (!u32_max(cands, ne)) {
    free(cands);
    return 0;
  }
  unmark_boundary(m, 1, cands);
  if (!u32_max(cands, ne)) {
    free(cands);
    return 0;
  }
  make_swap_tables();
  double* q;
  u32* rsz;
  swap_qualities(m, cands, &q, &rsz);
  if (!u32_max(cands, ne)) {
    free(cands);
    free(q);
    free(rsz);
    return 0;
  }
  u32* indset = find_indset(m, 1, cands, q);
  free(cands);
  free(q);
  u32 total = u32_sum(indset, ne);
  u32 nv = mesh_count(m, 0);
  orc_mesh* mo = mesh_new(m->dim, m->full);
  mesh_set_ents(mo, 0, nv, 0);
  for (u32 t = 0; t < m->tags[0].n; ++t) { /* copy_tags (tag.cpp) */
    otag* tg = m->tags[0].at[t];
    size_t bytes = tag_esz(tg->type) * (size_t) nv * tg->ncomps;
    void* d = xmalloc(bytes);
    memcpy(d, tg->data, bytes);
    tag_add(&mo->tags[0], tg->type, tg->name, tg->ncomps, d);
  }
  if (!m->full)
    swap_ents(m, mo, m->dim, indset, rsz);
  else
    for (u32 d = 1; d <= m->dim; ++d)
      swap_ents(m, mo, d, indset, rsz);
  free(indset);
  free(rsz);
  say("swapped", total, 1);
  mesh_overwrite(m, mo);
  return 1;
}

/* swap_slivers (swap.cpp:138-154) */
unsigned orc_swap_slivers(orc_mesh* m, double good_qual, unsigned nlayers)
{
  if (m->dim != 3)
    die("swap_slivers needs a 3D mesh");
  u32* sl = mark_slivers(m, good_qual, nlayers);
  u32* c = mark_down(m, m->dim, 1, sl);
  free(sl);
  return swap_common(m, c);
}

/* ------------------------------------------------------------- adapt
 * adapt.cpp:19-94 */
typedef struct {
  u32 ops, max_ops;
  int aborted;
  double npass, nelems, tpass;
} adapt_ctx;

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

static void adapt_summary(orc_mesh* m) /* adapt.cpp:22-34 */
{
  u32 nel = mesh_count(m, m->dim);
  double mq = mesh_min_quality(m);
  u32 ne = mesh_count(m, 1);
  double* s = measure_edges(m);
  double lo = orc_doubles_min(s, ne), hi = orc_doubles_max(s, ne);
  free(s);
  if (g_verbose)
    printf("%10u elements, min quality %.0f%%, metric range %.2f - %.2f\n",
        nel, mq * 100.0, lo, hi);
}

static int incr_op(orc_mesh* m, adapt_ctx* c) /* adapt.cpp:36-45 */
{
  if (c->ops > c->max_ops) {
    c->aborted = 1;
    return 0;
  }
  ++c->ops;
  adapt_summary(m);
  return 1;
}

typedef enum { P_REFINE, P_COARSEN, P_SWAP, P_COARSEN_SLIVERS } pass_kind;

static u32 run_pass(orc_mesh* m, adapt_ctx* c, pass_kind k, double a, double b, u32 n)
{
  double tin = (double) mesh_count(m, m->dim);
  double t0 = now_s();
  u32 ok = 0;
  switch (k) {
    case P_REFINE: ok = orc_refine_by_size(m, a); break;
    case P_COARSEN: ok = orc_coarsen_by_size(m, a, b); break;
    case P_SWAP: ok = orc_swap_slivers(m, a, n); break;
    case P_COARSEN_SLIVERS: ok = orc_coarsen_slivers(m, a, n); break;
  }
  double t1 = now_s();
  if (ok) {
    c->npass += 1;
    c->nelems += tin;
    c->tpass += t1 - t0;
  }
  return ok;
}

static int adapt_impl(orc_mesh* m, double size_floor, double good_qual,
    u32 nlayers, u32 max_ops, adapt_ctx* c)
{
  memset(c, 0, sizeof(*c));
  c->max_ops = max_ops;
  adapt_summary(m);
  /* satisfy_size (adapt.cpp:47-56) */
  double qf = mesh_min_quality(m);
  if (good_qual < qf)
    qf = good_qual;
  while (run_pass(m, c, P_REFINE, qf, 0, 0))
    if (!incr_op(m, c))
      return -1;
  while (run_pass(m, c, P_COARSEN, qf, size_floor, 0))
    if (!incr_op(m, c))
      return -1;
  /* satisfy_shape (adapt.cpp:58-80) */
  while (1) {
    double prev = mesh_min_quality(m);
    if (prev >= good_qual)
      break;
    if (m->dim == 3 && run_pass(m, c, P_SWAP, good_qual, 0, nlayers)) {
      if (!incr_op(m, c))
        return -1;
      continue;
    }
    if (run_pass(m, c, P_COARSEN_SLIVERS, good_qual, 0, nlayers)) {
      if (!incr_op(m, c))
        return -1;
      continue;
    }
    if (g_verbose)
      printf("warning: shape quality could not be satisfied\n");
    break;
  }
  return c->ops > 0;
}

/* mesh_adapt (adapt.cpp:82-94); returns -1 where the reference aborts
 * after max_ops operations */
int orc_adapt(orc_mesh* m, double size_floor, double good_qual, unsigned nlayers,
    unsigned max_ops)
{
  adapt_ctx c;
  return adapt_impl(m, size_floor, good_qual, nlayers, max_ops, &c);
}

int orc_adapt_timed(orc_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops, double* stats)
{
  adapt_ctx c;
  double t0 = now_s();
  int r = adapt_impl(m, size_floor, good_qual, nlayers, max_ops, &c);
  stats[0] = c.npass;
  stats[1] = c.nelems;
  stats[2] = c.tpass;
  stats[3] = now_s() - t0;
  return r;
}

/* mesh_warp_to_limit (warp_to_limit.cpp:14-81): halve the displacement
 * until the (serial) minimum quality reaches the floor */
unsigned orc_warp_to_limit(orc_mesh* m, double qual_floor)
{
  u32 dim = m->dim;
  u32 nv = mesh_count(m, 0), ne = mesh_count(m, dim);
  u32 const* ev = ask_down(m, dim, 0);
  double const* x = coords_of(m);
  otag* wt = find_tag(m, 0, "warp");
  if (!wt)
    die("no warp field");
  double const* w = (double const*) wt->data;
  double* xo = NEW(double, 3 * nv);
  double f = 1;
  int hit = 0;
  for (u32 i = 0; i < 3 * nv; ++i)
    xo[i] = f * w[i] + x[i];
  while (1) {
    double* q = element_qualities(dim, ne, ev, xo, metric_of(m));
    double mq = orc_doubles_min(q, ne);
    free(q);
    if (!(mq < qual_floor))
      break;
    if (g_verbose)
      printf("quality %f exceeds limit\n", mq);
    hit = 1;
    f /= 2;
    for (u32 i = 0; i < 3 * nv; ++i)
      xo[i] = f * w[i] + x[i];
  }
  double* wo = NEW(double, 3 * nv);
  for (u32 i = 0; i < 3 * nv; ++i)
    wo[i] = -f * w[i] + w[i];
  tag_remove(&m->tags[0], "coordinates");
  tag_remove(&m->tags[0], "warp");
  tag_add(&m->tags[0], T_F64, "coordinates", 3, xo);
  tag_add(&m->tags[0], T_F64, "warp", 3, wo);
  return !hit;
}

/* ------------------------------------------------------ flat C ABI */
orc_mesh* orc_new_mesh(unsigned dim, unsigned nverts, double const* coords,
    unsigned nelems, unsigned const* conn)
{
  orc_mesh* m = mesh_new(dim, 0);
  mesh_set_ents(m, 0, nverts, 0);
  mesh_set_ents(m, dim, nelems, u32_copy(conn, nelems * (dim + 1)));
  tag_add(&m->tags[0], T_F64, "coordinates", 3, f64_copy(coords, 3 * nverts));
  return m;
}

unsigned orc_dim(orc_mesh* m) { return m->dim; }
unsigned orc_count(orc_mesh* m, unsigned dim) { return mesh_count(m, dim); }

void orc_get_down(orc_mesh* m, unsigned h, unsigned l, unsigned* out)
{
  u32 n = mesh_count(m, h) * down_deg[h][l];
  memcpy(out, ask_down(m, h, l), sizeof(u32) * n);
}

unsigned orc_up_nuses(orc_mesh* m, unsigned l, unsigned h)
{
  oup const* u = ask_up(m, l, h);
  return u->offsets[mesh_count(m, l)];
}

void orc_get_up(orc_mesh* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj, unsigned* dirs)
{
  oup const* u = ask_up(m, l, h);
  u32 n = mesh_count(m, l);
  memcpy(offs, u->offsets, sizeof(u32) * (n + 1));
  memcpy(adj, u->adj, sizeof(u32) * u->offsets[n]);
  memcpy(dirs, u->dirs, sizeof(u32) * u->offsets[n]);
}

unsigned orc_star_nadj(orc_mesh* m, unsigned l, unsigned h)
{
  ograph const* g = ask_star(m, l, h);
  return g->offsets[mesh_count(m, l)];
}

void orc_get_star(orc_mesh* m, unsigned l, unsigned h, unsigned* offs, unsigned* adj)
{
  ograph const* g = ask_star(m, l, h);
  u32 n = mesh_count(m, l);
  memcpy(offs, g->offsets, sizeof(u32) * (n + 1));
  memcpy(adj, g->adj, sizeof(u32) * g->offsets[n]);
}

void orc_get_dual(orc_mesh* m, unsigned* out)
{
  u32 d = m->dim;
  memcpy(out, ask_dual(m), sizeof(u32) * mesh_count(m, d) * down_deg[d][d - 1]);
}

int orc_get_tag(orc_mesh* m, unsigned dim, char const* name, int want_f64, void* out)
{
  otag* t = find_tag(m, dim, name);
  if (!t)
    return -1;
  if ((t->type == T_F64) != (want_f64 != 0))
    return -1;
  if (out)
    memcpy(out, t->data, tag_esz(t->type) * (size_t) mesh_count(m, dim) * t->ncomps);
  return (int) t->ncomps;
}

unsigned orc_ntags(orc_mesh* m, unsigned dim) { return m->tags[dim].n; }
char const* orc_tag_name(orc_mesh* m, unsigned dim, unsigned i)
{
  return m->tags[dim].at[i]->name;
}

void orc_add_field(orc_mesh* m, unsigned dim, char const* name, unsigned ncomps,
    double const* data)
{
  u32 n = mesh_count(m, dim) * ncomps;
  tag_add(&m->tags[dim], T_F64, name, ncomps, f64_copy(data, n));
}

/* the fit of coarsen_fit on its own: rank of the npts x 4 row-major a,
 * and the least-squares solution when the rank is 4 */
unsigned orc_qr_fit(unsigned npts, double const* a, double const* b, double* x)
{
  double aa[MAX_PTS][4], q[MAX_PTS][MAX_PTS], r[MAX_PTS][4];
  if (npts < 4 || npts > MAX_PTS)
    die("orc_qr_fit: 4 to 16 points");
  for (u32 i = 0; i < npts; ++i)
    for (u32 j = 0; j < 4; ++j)
      aa[i][j] = a[i * 4 + j];
  u32 rank = qr_decomp2(aa, q, r, npts);
  if (rank == 4)
    qr_solve2(q, r, b, npts, x);
  return rank;
}

/* osh_new_field with a transfer type (omega_h.cpp:613-622) */
void orc_add_field_tt(orc_mesh* m, unsigned dim, char const* name, unsigned ncomps,
    double const* data, int transfer)
{
  u32 n = mesh_count(m, dim) * ncomps;
  tag_add2(&m->tags[dim], T_F64, name, ncomps, transfer, f64_copy(data, n));
}

void orc_free_field(orc_mesh* m, unsigned dim, char const* name)
{
  tag_remove(&m->tags[dim], name);
}

double orc_min_quality(orc_mesh* m) { return mesh_min_quality(m); }

void orc_qualities(orc_mesh* m, double* out)
{
  double* q = mesh_qualities(m);
  memcpy(out, q, sizeof(double) * mesh_count(m, m->dim));
  free(q);
}

void orc_measure_edges(orc_mesh* m, double* out)
{
  double* s = measure_edges(m);
  memcpy(out, s, sizeof(double) * mesh_count(m, 1));
  free(s);
}

/* metric from a Hessian: cyclic Jacobi eigendecomposition in FP64 and
 * eigenvalue clamping; the extension of size_from_hessian.cpp:12-40 that the
 * device code implements, in the same operation order */
static void jacobi_eigen3(double const* h6, double* lam, double V[3][3])
{
  double A[3][3] = {{h6[0], h6[1], h6[2]}, {h6[1], h6[3], h6[4]}, {h6[2], h6[4], h6[5]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      V[i][j] = (i == j) ? 1.0 : 0.0;
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
    if (off <= 1e-40 * dia || off == 0.0)
      break;
    for (int r = 0; r < 3; ++r) {
      int p = P[r], q = Q[r];
      double apq = A[p][q];
      if (apq == 0.0)
        continue;
      double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
      double at = theta < 0 ? -theta : theta;
      double t = 1.0 / (at + sqrt(theta * theta + 1.0));
      if (theta < 0)
        t = -t;
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

void orc_eigen3(double const* h6, double* lam, double* V9)
{
  double V[3][3];
  jacobi_eigen3(h6, lam, V);
  memcpy(V9, V, sizeof(V));
}

void orc_metric_from_hessian(orc_mesh* m, char const* hname, double eps, double hmin,
    double hmax, char const* out_name)
{
  otag* t = find_tag(m, 0, hname);
  if (!t || t->type != T_F64 || t->ncomps != 6)
    die("hessian field must have 6 FP64 components");
  u32 nv = mesh_count(m, 0);
  double const* h = (double const*) t->data;
  double* out = NEW(double, 6 * nv);
  double lo = 1.0 / (hmax * hmax), hi = 1.0 / (hmin * hmin);
  int const I[6] = {0, 0, 0, 1, 1, 2}, J[6] = {0, 1, 2, 1, 2, 2};
  for (u32 v = 0; v < nv; ++v) {
    double lam[3], V[3][3], l[3];
    jacobi_eigen3(h + 6 * v, lam, V);
    for (int k = 0; k < 3; ++k) {
      double a = lam[k] < 0 ? -lam[k] : lam[k];
      a = a / eps;
      if (a < lo)
        a = lo;
      if (a > hi)
        a = hi;
      l[k] = a;
    }
    for (int c = 0; c < 6; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k)
        s += (V[I[c]][k] * l[k]) * V[J[c]][k];
      out[6 * v + c] = s;
    }
  }
  if (find_tag(m, 0, out_name))
    tag_remove(&m->tags[0], out_name);
  tag_add(&m->tags[0], T_F64, out_name, 6, out);
}

double orc_metric_quality(unsigned dim, double const* x, double const* m6, int signed_z)
{
  return metric_quality(dim, (double const (*)[3]) x, (double const (*)[6]) m6, signed_z);
}

double orc_metric_length(double const* a6, double const* b6, double const* x0, double const* x1)
{
  double e[3];
  sub3(x1, x0, e);
  return sqrt(metric_len2(a6, b6, e));
}

/* test access to the candidate-quality loops: refine_qualities on edges
 * (refine_qualities.cpp:84-133) and coarsen_qualities
 * (coarsen_qualities.cpp:81-114); cands / codes are updated in place as
 * the passes update them, out gets the goodness slots (ne, or 2 ne) */
void orc_refine_qualities(orc_mesh* m, unsigned* cands, double qual_floor,
    int require_better, double* out)
{
  double* q = refine_qualities(m, 1, cands, qual_floor, require_better);
  memcpy(out, q, sizeof(double) * mesh_count(m, 1));
  free(q);
}

void orc_coarsen_qualities(orc_mesh* m, unsigned* codes, double qual_floor,
    int require_better, double* out)
{
  double* q = coarsen_qualities(m, codes, qual_floor, require_better);
  memcpy(out, q, sizeof(double) * 2 * mesh_count(m, 1));
  free(q);
}

/* ------------------------------------------------ size-field producers
 * The warp loop's size field (test_warp.cpp:70-83): element gradients of a
 * vertex field, recovered to vertices by volume weighting, twice, then a
 * size from the Hessian's norm; and the identity size field. */

/* element_jacobian (jacobian.cpp:7-17): column i = edge from vertex 0 to
 * vertex i + 1 */
static void element_jacobian(u32 nv, double (*x)[3], double jac[3][3])
{
  for (u32 i = 0; i < 3; ++i)
    for (u32 j = 0; j < 3; ++j)
      jac[i][j] = 0;
  for (u32 i = 0; i < nv; ++i)
    for (u32 j = 0; j < 3; ++j)
      jac[j][i] = x[i + 1][j] - x[0][j];
}

/* invert_jacobian (jacobian.cpp:19-66, algebra.hpp:143-196) for 2 and 3 */
static void invert_jacobian(u32 dim, double in[3][3], double out[3][3])
{
  if (dim == 3) {
    double tmp[3][3];
    double c[3];
    cross(in[0], in[1], c);
    double d = dot3(in[2], c);
    cross(in[1], in[2], tmp[0]);
    cross(in[2], in[0], tmp[1]);
    cross(in[0], in[1], tmp[2]);
    double s = 1.0 / d;
    for (u32 i = 0; i < 3; ++i)
      for (u32 j = 0; j < 3; ++j)
        out[i][j] = tmp[j][i] * s;
    return;
  }
  double nrm[2][2];
  for (u32 i = 0; i < 2; ++i)
    for (u32 j = 0; j < 2; ++j) {
      nrm[i][j] = 0;
      for (u32 k = 0; k < 3; ++k)
        nrm[i][j] += in[i][k] * in[j][k];
    }
  double det = nrm[0][0] * nrm[1][1] - nrm[0][1] * nrm[1][0];
  double inv[2][2];
  inv[0][0] = nrm[1][1] / det;
  inv[0][1] = -nrm[0][1] / det;
  inv[1][0] = -nrm[1][0] / det;
  inv[1][1] = nrm[0][0] / det;
  for (u32 i = 0; i < 3; ++i)
    for (u32 j = 0; j < 3; ++j)
      out[i][j] = 0;
  for (u32 i = 0; i < 3; ++i)
    for (u32 j = 0; j < 2; ++j)
      for (u32 k = 0; k < 2; ++k)
        out[i][j] += in[k][i] * inv[k][j];
}

/* mesh_element_gradients (element_gradients.cpp:15-91): element tag
 * "grad_<name>" with 3 components per input component */
void orc_element_gradients(orc_mesh* m, char const* name)
{
  u32 dim = m->dim, n = mesh_count(m, dim), vpe = dim + 1;
  otag* t = find_tag(m, 0, name);
  if (!t || t->type != T_F64)
    die("element_gradients: no such vertex field");
  u32 nc = t->ncomps, nco = 3 * nc;
  double const* f = (double const*) t->data;
  double const* x = coords_of(m);
  u32 const* ev = ask_down(m, dim, 0);
  double* out = NEW(double, (size_t) n * nco);
  for (u32 i = 0; i < n; ++i) {
    double ex[4][3];
    double const* ec[4];
    for (u32 j = 0; j < vpe; ++j) {
      u32 v = ev[i * vpe + j];
      ec[j] = f + (size_t) v * nc;
      memcpy(ex[j], x + 3 * (size_t) v, sizeof(double) * 3);
    }
    double jac[3][3], jaci[3][3];
    element_jacobian(dim, ex, jac);
    invert_jacobian(dim, jac, jaci);
    double* g = out + (size_t) i * nco;
    for (u32 j = 0; j < nco; ++j)
      g[j] = 0;
    for (u32 j = 0; j < dim; ++j)
      for (u32 k = 0; k < nc; ++k)
        for (u32 l = 0; l < 3; ++l)
          g[k * 3 + l] += jaci[j][l] * (ec[j + 1][k] - ec[0][k]);
  }
  char gname[80];
  snprintf(gname, sizeof(gname), "grad_%s", name);
  tag_add(&m->tags[dim], T_F64, gname, nco, out);
}

/* mesh_recover_by_volume (recover_by_volume.cpp:10-68): the vertex field
 * name = the size-weighted mean of the element field name around it */
void orc_recover_by_volume(orc_mesh* m, char const* name)
{
  u32 dim = m->dim, nv = mesh_count(m, 0);
  otag* t = find_tag(m, dim, name);
  if (!t || t->type != T_F64)
    die("recover_by_volume: no such element field");
  u32 nc = t->ncomps;
  double const* f = (double const*) t->data;
  double* sizes = element_sizes(m);
  oup const* up = ask_up(m, 0, dim);
  double* out = NEW(double, (size_t) nv * nc);
  for (u32 i = 0; i < nv; ++i) {
    double* o = out + (size_t) i * nc;
    for (u32 j = 0; j < nc; ++j)
      o[j] = 0;
    double ssum = 0;
    for (u32 j = up->offsets[i]; j < up->offsets[i + 1]; ++j) {
      u32 el = up->adj[j];
      double es = sizes[el];
      ssum += es;
      for (u32 k = 0; k < nc; ++k)
        o[k] += es * f[(size_t) el * nc + k];
    }
    for (u32 j = 0; j < nc; ++j)
      o[j] /= ssum;
  }
  free(sizes);
  tag_add(&m->tags[0], T_F64, name, nc, out);
}

/* mesh_size_from_hessian (size_from_hessian.cpp:11-72): per vertex the
 * weighted sum of the Frobenius norms of the 3x3 Hessians (one per
 * solution component), then max_h - that, floored at min_h, as the vertex
 * field "adapt_size" */
void orc_size_from_hessian(orc_mesh* m, char const* hname, double const* weights, double min_h,
    double max_h)
{
  u32 nv = mesh_count(m, 0);
  otag* t = find_tag(m, 0, hname);
  if (!t || t->type != T_F64 || t->ncomps % 9)
    die("size_from_hessian: needs a vertex field of 9 components per solution component");
  if (!(max_h > min_h) || !(min_h > 0))
    die("size_from_hessian: needs 0 < min_h < max_h");
  u32 nh = t->ncomps, ns = nh / 9;
  double const* h = (double const*) t->data;
  double* out = NEW(double, nv);
  for (u32 i = 0; i < nv; ++i) {
    double const* hv = h + (size_t) i * nh;
    double total = 0;
    for (u32 j = 0; j < ns; ++j) {
      double d = 0;
      for (u32 k = 0; k < 9; ++k)
        d += hv[k] * hv[k];
      double nrm = sqrt(d);
      double w = weights ? weights[j] : 1;
      if (!(w >= 0))
        die("size_from_hessian: negative weight");
      total += w * nrm;
      hv += 9;
    }
    out[i] = total;
  }
  for (u32 i = 0; i < nv; ++i) {
    out[i] = max_h - out[i];
    if (out[i] < min_h)
      out[i] = min_h;
  }
  tag_add(&m->tags[0], T_F64, "adapt_size", 1, out);
}

/* mesh_identity_size_field (size.cpp:90-135): the longest edge at each
 * vertex */
void orc_identity_size(orc_mesh* m, char const* name)
{
  u32 nv = mesh_count(m, 0);
  ograph const* g = ask_star(m, 0, 1);
  double const* x = coords_of(m);
  double* out = NEW(double, nv);
  for (u32 i = 0; i < nv; ++i) {
    double mx = 0;
    for (u32 j = g->offsets[i]; j < g->offsets[i + 1]; ++j) {
      double l = sqrt(vsqdist(x + 3 * (size_t) g->adj[j], x + 3 * (size_t) i));
      if (l > mx)
        mx = l;
    }
    out[i] = mx;
  }
  tag_add(&m->tags[0], T_F64, name, 1, out);
}
