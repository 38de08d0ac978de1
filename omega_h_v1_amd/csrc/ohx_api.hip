// C ABI: the public osh_* drop-in, the ohx_mesh_* seam and the ohx_*
// kernel entry points.  Exceptions never cross the ABI: the seam and kernel
// layers record ohx_last_error() and return a sentinel, the osh_* layer
// aborts like the reference's asserts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <chrono>
#include <mutex>
#include <functional>
#include <condition_variable>
#include <thread>
#include <memory>
#include <string>
#include <type_traits>

#include "../../include/ohx.h"
#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_prims.hpp"
#include "ohx_tables.hpp"

namespace ohx {
std::string const& last_error();
void set_last_error(std::string const& s);
void prof_enable(bool on);
void prof_reset();
bool prof_get(char const* name, double* ms, u64* launches, double* units);
std::vector<std::string> prof_names();
}  // namespace ohx

using namespace ohx;

struct ohx_mesh {
  Mesh m;
  // per-dimension caches handed out through osh_* getters
  Buf<u32> own_rank[4], own_id[4];
  Buf<unsigned long> global[4];
  std::vector<std::string> field_names;
  std::unique_ptr<BatchCtx> batch;  // set on the mesh of a batch of meshes
  ohx_mesh(u32 d, bool f) : m(d, f) {}
};

static thread_local std::string t_err;
static thread_local bool t_has_err = false;

namespace ohx {
// for entry points defined in other sources
void api_set_error(char const* what) {
  t_err = what;
  t_has_err = true;
}
void noalloc_setup_current_device();  // ohx_noalloc.hip
}  // namespace ohx

template <class F, class R>
static R guard(R bad, F&& f) {
  try {
    return f();
  } catch (std::exception const& e) {
    t_err = e.what();
    t_has_err = true;
    return bad;
  }
}

template <class F>
static void guard_void(F&& f) {
  try {
    f();
  } catch (std::exception const& e) {
    t_err = e.what();
    t_has_err = true;
  }
}

// the public layer mirrors the reference's assert/abort behaviour, and
// returns only once the library's stream has finished: the caller reads the
// returned arrays and fills the new ones from its own stream (the
// reference's arrays are ready when its calls return)
template <class F>
static auto osh_guard(F&& f) -> decltype(f()) {
  try {
    if constexpr (std::is_void_v<decltype(f())>) {
      f();
      sync();
    } else {
      auto r = f();
      sync();
      return r;
    }
  } catch (std::exception const& e) {
    std::fprintf(stderr, "omega_h (MI355X): %s\n", e.what());
    std::abort();
  }
}

static ohx_mesh* M(osh_t p) { return reinterpret_cast<ohx_mesh*>(p); }

// ------------------------------------------------------- Kuhn generator
static ohx_mesh* new_box(u32 dim, u32 n) {
  OHX_REQUIRE(dim == 2 || dim == 3, "box dimension must be 2 or 3");
  OHX_REQUIRE(n >= 1, "box needs at least one cell per axis");
  u32 np = n + 1;
  u64 nv64 = (dim == 3) ? (u64) np * np * np : (u64) np * np;
  u64 nc64 = (dim == 3) ? (u64) n * n * n : (u64) n * n;
  OHX_REQUIRE(nc64 * (dim == 3 ? 6 : 2) * (dim + 1) < (1ull << 32), "box too large for 32-bit indices");
  u32 nv = (u32) nv64, nc = (u32) nc64;
  u32 per = (dim == 3) ? 6 : 2;
  auto* mm = new ohx_mesh(dim, false);
  Mesh& m = mm->m;
  Buf<double> x((size_t) 3 * nv);
  double* xp = x.get();
  double dn = (double) n;
  parallel_for(nv, OHX_LAMBDA(u32 v) {
    u32 i = v % np, j = (v / np) % np, k = v / (np * np);
    xp[3 * (size_t) v + 0] = (double) i / dn;
    xp[3 * (size_t) v + 1] = (double) j / dn;
    xp[3 * (size_t) v + 2] = (dim == 3) ? (double) k / dn : 0.0;
  });
  Buf<u32> conn((size_t) nc * per * (dim + 1));
  u32* cp = conn.get();
  if (dim == 3) {
    parallel_for(nc, OHX_LAMBDA(u32 c) {
      // corners and split of box_3d (tables.cpp:29-47)
      u32 const cx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cy[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
      u32 const t[24] = {0, 1, 2, 6, 2, 3, 0, 6, 0, 3, 7, 6, 7, 4, 0, 6, 0, 4, 5, 6, 5, 1, 0, 6};
      u32 ci = c % n, cj = (c / n) % n, ck = c / (n * n);
      u32 vid[8];
      for (u32 a = 0; a < 8; ++a) vid[a] = (ci + cx[a]) + np * ((cj + cy[a]) + np * (ck + cz[a]));
      for (u32 a = 0; a < 24; ++a) cp[(size_t) c * 24 + a] = vid[t[a]];
    });
  } else {
    parallel_for(nc, OHX_LAMBDA(u32 c) {
      u32 const cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};  // box_2d (tables.cpp:15-27)
      u32 const t[6] = {0, 1, 2, 2, 3, 0};
      u32 ci = c % n, cj = c / n;
      u32 vid[4];
      for (u32 a = 0; a < 4; ++a) vid[a] = (ci + cx[a]) + np * (cj + cy[a]);
      for (u32 a = 0; a < 6; ++a) cp[(size_t) c * 6 + a] = vid[t[a]];
    });
  }
  set_ents(m, 0, nv, Buf<u32>());
  set_ents(m, dim, nc * per, std::move(conn));
  add_tag_f64(m, 0, "coordinates", 3, std::move(x));
  return mm;
}

static ohx_mesh* clone_mesh(ohx_mesh* src) {
  Mesh& s = src->m;
  auto* mm = new ohx_mesh(s.dim, s.full);
  Mesh& m = mm->m;
  for (u32 d = 0; d <= s.dim; ++d) {
    if (!has_dim(s, d)) continue;
    u32 n = s.counts[d];
    if (d == 0)
      set_ents(m, 0, n, Buf<u32>());
    else
      set_ents(m, d, n, copy(s.down[d][0].get(), n * (d + 1)));
    for (auto& tp : s.tags[d]) {
      Tag& t = *tp;
      if (t.type == TagType::F64)
        add_tag_f64(m, d, t.name.c_str(), t.ncomps, copy_f64(t.f.get(), n * t.ncomps), t.transfer);
      else
        add_tag_u32(m, d, t.name.c_str(), t.ncomps, copy(t.u.get(), n * t.ncomps), t.transfer);
    }
  }
  return mm;
}

extern "C" {

char const* ohx_last_error(void) { return t_has_err ? t_err.c_str() : nullptr; }
void ohx_clear_error(void) {
  t_has_err = false;
  t_err.clear();
}

// ------------------------------------------------------------ kernel ABI
unsigned ohx_element_qualities_f64(unsigned dim, unsigned n, unsigned const* ev, double const* x,
                                   double* out, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    OHX_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    StreamScope sc((hipStream_t) s);
    Buf<double> q = element_qualities(dim, n, ev, x);
    d2d(out, q.get(), n);
    sync();
    return 0u;
  });
}

double ohx_min_quality_f64(unsigned dim, unsigned n, unsigned const* ev, double const* x, void* s) {
  return guard(std::numeric_limits<double>::quiet_NaN(), [&]() -> double {
    OHX_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    StreamScope sc((hipStream_t) s);
    return min_element_quality(dim, n, ev, x);
  });
}

unsigned ohx_measure_edges_f64(unsigned ne, unsigned const* ev, double const* x, double const* h,
                               double* out, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    parallel_for(ne, OHX_LAMBDA(u32 i) {
      u32 a = ev[2 * i], b = ev[2 * i + 1];
      double len = sqrt(sqdist(ld3(x + 3 * (size_t) b), ld3(x + 3 * (size_t) a)));
      double want = (h[a] + h[b]) / 2;
      out[i] = len / want;
    });
    sync();
    return 0u;
  });
}

unsigned ohx_exscan_u32(unsigned const* a, unsigned n, unsigned* out, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    exscan(a, n, out);
    sync();
    return 0u;
  });
}

unsigned ohx_up_from_down(unsigned hd, unsigned ld, unsigned nh, unsigned nl, unsigned const* lh,
                          unsigned* offs, unsigned* adj, unsigned* dirs, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    OHX_REQUIRE(hd <= 3 && ld < hd, "bad dimensions");
    StreamScope sc((hipStream_t) s);
    Up u;
    up_from_down(hd, ld, nh, nl, lh, u);
    u32 nuses = nh * h_down_deg[hd][ld];
    d2d(offs, u.offsets.get(), (size_t) nl + 1);
    d2d(adj, u.adj.get(), nuses);
    d2d(dirs, u.dirs.get(), nuses);
    sync();
    return 0u;
  });
}

unsigned ohx_find_indset(unsigned n, unsigned const* offs, unsigned const* adj, unsigned const* filter,
                         double const* good, unsigned* state, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    // a one-dimensional "mesh" whose vertex star is the given graph
    Mesh m(1, true);
    set_ents(m, 0, n, Buf<u32>());
    auto g = std::make_unique<Graph>();
    g->offsets = copy(offs, n + 1);
    u32 nadj = read_at(offs, n);
    g->adj = copy(adj, nadj);
    m.star[0][1] = std::move(g);
    Buf<u32> st = find_indset(m, 0, filter, good);
    d2d(state, st.get(), n);
    sync();
    return 0u;
  });
}

unsigned ohx_fp64_selftest(unsigned n, double const* a, double const* b, double* so, double* dq, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    parallel_for(n, OHX_LAMBDA(u32 i) {
      so[i] = sqrt(a[i]);
      dq[i] = a[i] / b[i];
    });
    sync();
    return 0u;
  });
}

unsigned ohx_eigen3_f64(unsigned n, double const* h6, double* lam, double* V9, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    parallel_for(n, OHX_LAMBDA(u32 i) {
      double V[3][3];
      jacobi_eigen3(h6 + 6 * (size_t) i, lam + 3 * (size_t) i, V);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) V9[9 * (size_t) i + 3 * a + b] = V[a][b];
    });
    sync();
    return 0u;
  });
}

unsigned ohx_metric_from_hessian_f64(unsigned n, double const* h6, double eps, double hmin, double hmax,
                                     double* m6, void* s) {
  return guard(OHX_ERROR, [&]() -> unsigned {
    StreamScope sc((hipStream_t) s);
    parallel_for(n, OHX_LAMBDA(u32 i) { metric_from_hessian(h6 + 6 * (size_t) i, eps, hmin, hmax, m6 + 6 * (size_t) i); });
    sync();
    return 0u;
  });
}

// ------------------------------------------------------------ mesh seam
ohx_mesh* ohx_mesh_new(unsigned dim, unsigned nv, double const* x, unsigned ne, unsigned const* conn) {
  return guard((ohx_mesh*) nullptr, [&]() -> ohx_mesh* {
    OHX_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    auto* mm = new ohx_mesh(dim, false);
    set_ents(mm->m, 0, nv, Buf<u32>());
    set_ents(mm->m, dim, ne, to_device(conn, (size_t) ne * (dim + 1)));
    add_tag_f64(mm->m, 0, "coordinates", 3, to_device(x, (size_t) 3 * nv));
    return mm;
  });
}

ohx_mesh* ohx_mesh_new_box(unsigned dim, unsigned n) {
  return guard((ohx_mesh*) nullptr, [&]() { return new_box(dim, n); });
}

ohx_mesh* ohx_mesh_clone(ohx_mesh* m) {
  return guard((ohx_mesh*) nullptr, [&]() { return clone_mesh(m); });
}

void ohx_mesh_free(ohx_mesh* m) {
  guard_void([&]() { delete m; });
}

void ohx_mesh_derive_model_full(ohx_mesh* m) {
  guard_void([&]() {
    derive_model(m->m, REF_PI / 4);
    m->m.full = true;
  });
}

void ohx_mesh_set_full(ohx_mesh* m) { m->m.full = true; }

unsigned ohx_mesh_count(ohx_mesh* m, unsigned d) {
  return guard(OHX_ERROR, [&]() { return count(m->m, d); });
}

void ohx_mesh_get_down(ohx_mesh* m, unsigned h, unsigned l, unsigned* out) {
  guard_void([&]() {
    u32 n = count(m->m, h) * h_down_deg[h][l];
    to_host(out, ask_down(m->m, h, l), n);
  });
}

unsigned ohx_mesh_up_nuses(ohx_mesh* m, unsigned l, unsigned h) {
  return guard(OHX_ERROR, [&]() {
    Up const& u = ask_up(m->m, l, h);
    return read_at(u.offsets.get(), count(m->m, l));
  });
}

void ohx_mesh_get_up(ohx_mesh* m, unsigned l, unsigned h, unsigned* offs, unsigned* adj, unsigned* dirs) {
  guard_void([&]() {
    Up const& u = ask_up(m->m, l, h);
    u32 n = count(m->m, l);
    to_host(offs, u.offsets.get(), n + 1);
    to_host(adj, u.adj.get(), offs[n]);
    to_host(dirs, u.dirs.get(), offs[n]);
  });
}

unsigned ohx_mesh_star_nadj(ohx_mesh* m, unsigned l, unsigned h) {
  return guard(OHX_ERROR, [&]() {
    Graph const& g = ask_star(m->m, l, h);
    return read_at(g.offsets.get(), count(m->m, l));
  });
}

void ohx_mesh_get_star(ohx_mesh* m, unsigned l, unsigned h, unsigned* offs, unsigned* adj) {
  guard_void([&]() {
    Graph const& g = ask_star(m->m, l, h);
    u32 n = count(m->m, l);
    to_host(offs, g.offsets.get(), n + 1);
    to_host(adj, g.adj.get(), offs[n]);
  });
}

void ohx_mesh_get_dual(ohx_mesh* m, unsigned* out) {
  guard_void([&]() {
    u32 d = m->m.dim;
    to_host(out, ask_dual(m->m), (size_t) count(m->m, d) * h_down_deg[d][d - 1]);
  });
}

int ohx_mesh_get_tag(ohx_mesh* m, unsigned d, char const* name, int want_f64, void* out) {
  return guard(-1, [&]() -> int {
    Tag* t = find_tag(m->m, d, name);
    if (!t) return -1;
    if ((t->type == TagType::F64) != (want_f64 != 0)) return -1;
    if (out) {
      size_t n = (size_t) count(m->m, d) * t->ncomps;
      if (t->type == TagType::F64)
        to_host((double*) out, t->f.get(), n);
      else
        to_host((u32*) out, t->u.get(), n);
    }
    return (int) t->ncomps;
  });
}

unsigned ohx_mesh_ntags(ohx_mesh* m, unsigned d) { return (unsigned) m->m.tags[d].size(); }

char const* ohx_mesh_tag_name(ohx_mesh* m, unsigned d, unsigned i) {
  return m->m.tags[d][i]->name.c_str();
}

void ohx_mesh_add_field(ohx_mesh* m, unsigned d, char const* name, unsigned nc, double const* data) {
  guard_void([&]() {
    size_t n = (size_t) count(m->m, d) * nc;
    add_tag_f64(m->m, d, name, nc, to_device(data, n));
  });
}

void ohx_mesh_add_field_tt(ohx_mesh* m, unsigned d, char const* name, unsigned nc, double const* data,
                           int transfer) {
  guard_void([&]() {
    OHX_REQUIRE(transfer >= 0 && transfer <= 3, "transfer must be an enum osh_transfer value");
    size_t n = (size_t) count(m->m, d) * nc;
    add_tag_f64(m->m, d, name, nc, to_device(data, n), transfer);
  });
}

void ohx_mesh_element_gradients(ohx_mesh* m, char const* name) {
  guard_void([&]() { element_gradients(m->m, name); });
}

void ohx_mesh_recover_by_volume(ohx_mesh* m, char const* name) {
  guard_void([&]() { recover_by_volume(m->m, name); });
}

void ohx_mesh_size_from_hessian(ohx_mesh* m, char const* hess_name, double const* weights, double min_h,
                                double max_h) {
  guard_void([&]() { size_from_hessian(m->m, hess_name, weights, min_h, max_h); });
}

void ohx_mesh_identity_size(ohx_mesh* m, char const* name) {
  guard_void([&]() { identity_size(m->m, name); });
}

void ohx_mesh_free_field(ohx_mesh* m, unsigned d, char const* name) {
  guard_void([&]() { remove_tag(m->m, d, name); });
}

void ohx_mesh_set_uniform_field(ohx_mesh* m, char const* name, double v) {
  guard_void([&]() {
    if (find_tag(m->m, 0, name)) remove_tag(m->m, 0, name);
    add_tag_f64(m->m, 0, name, 1, filled_f64(count(m->m, 0), v));
  });
}

void ohx_mesh_metric_from_hessian(ohx_mesh* m, char const* hname, double eps, double hmin, double hmax,
                                  char const* out_name) {
  guard_void([&]() {
    Tag* t = find_tag(m->m, 0, hname);
    OHX_REQUIRE(t && t->type == TagType::F64 && t->ncomps == 6, "hessian field must have 6 FP64 components");
    u32 nv = count(m->m, 0);
    Buf<double> out((size_t) 6 * nv);
    double const* h = t->f.get();
    double* o = out.get();
    parallel_for(nv, OHX_LAMBDA(u32 i) { metric_from_hessian(h + 6 * (size_t) i, eps, hmin, hmax, o + 6 * (size_t) i); });
    if (find_tag(m->m, 0, out_name)) remove_tag(m->m, 0, out_name);
    add_tag_f64(m->m, 0, out_name, 6, std::move(out));
  });
}

double ohx_mesh_min_quality(ohx_mesh* m) {
  return guard(std::numeric_limits<double>::quiet_NaN(), [&]() { return mesh_min_quality(m->m); });
}

void ohx_mesh_qualities(ohx_mesh* m, double* out) {
  guard_void([&]() {
    Buf<double> q = mesh_qualities(m->m);
    to_host(out, q.get(), count(m->m, m->m.dim));
  });
}

void ohx_mesh_measure_edges(ohx_mesh* m, double* out) {
  guard_void([&]() {
    Buf<double> s = measure_edges(m->m);
    to_host(out, s.get(), count(m->m, 1));
  });
}

void ohx_mesh_find_indset(ohx_mesh* m, unsigned ed, unsigned const* cand, double const* good, unsigned* out) {
  guard_void([&]() {
    u32 n = count(m->m, ed);
    Buf<u32> c = to_device(cand, n);
    Buf<double> g = to_device(good, n);
    Buf<u32> s = find_indset(m->m, ed, c.get(), g.get());
    to_host(out, s.get(), n);
  });
}

unsigned ohx_mesh_refine_by_size(ohx_mesh* m, double qf) {
  return guard(OHX_ERROR, [&]() { return refine_by_size(m->m, qf); });
}

void ohx_mesh_uniformly_refine(ohx_mesh* m) {
  guard_void([&]() { uniformly_refine(m->m); });
}

unsigned ohx_mesh_coarsen_by_size(ohx_mesh* m, double qf, double sf) {
  return guard(OHX_ERROR, [&]() { return coarsen_by_size(m->m, qf, sf); });
}

unsigned ohx_mesh_coarsen_slivers(ohx_mesh* m, double qf, unsigned nl) {
  return guard(OHX_ERROR, [&]() { return coarsen_slivers(m->m, qf, nl); });
}

unsigned ohx_mesh_swap_slivers(ohx_mesh* m, double gq, unsigned nl) {
  return guard(OHX_ERROR, [&]() { return swap_slivers(m->m, gq, nl); });
}

int ohx_mesh_adapt(ohx_mesh* m, double sf, double gq, unsigned nl, unsigned mo) {
  return guard(-2, [&]() { return adapt(m->m, sf, gq, nl, mo, nullptr); });
}

int ohx_mesh_adapt_timed(ohx_mesh* m, double sf, double gq, unsigned nl, unsigned mo, double* stats) {
  return guard(-2, [&]() {
    AdaptStats st;
    int r = adapt(m->m, sf, gq, nl, mo, &st);
    stats[0] = st.npass;
    stats[1] = st.nelems;
    stats[2] = st.tpass;
    stats[3] = st.ttotal;
    return r;
  });
}

ohx_mesh* ohx_mesh_batch_new(unsigned n, ohx_mesh* const* meshes) {
  return guard((ohx_mesh*) nullptr, [&]() {
    OHX_REQUIRE(n > 0 && meshes, "batch: no meshes");
    std::vector<Mesh*> parts;
    for (unsigned i = 0; i < n; ++i) parts.push_back(&meshes[i]->m);
    auto* mm = new ohx_mesh(parts[0]->dim, true);
    mm->batch.reset(new BatchCtx);
    try {
      batch_union(parts, mm->m, *mm->batch);
      sync();
    } catch (...) {
      delete mm;
      throw;
    }
    return mm;
  });
}

unsigned ohx_mesh_batch_size(ohx_mesh* m) { return m->batch ? m->batch->ncomp : 0; }

void ohx_set_batch_retire_min(double elements) { g_batch_retire_min = elements; }
void ohx_set_batch_handoff(double fraction) { g_batch_handoff = fraction; }

ohx_mesh* ohx_mesh_batch_extract(ohx_mesh* m, unsigned c) {
  return guard((ohx_mesh*) nullptr, [&]() {
    OHX_REQUIRE(m->batch, "batch: not a batch mesh");
    auto* mm = new ohx_mesh(m->m.dim, m->m.full);
    try {
      batch_extract(m->m, *m->batch, c, mm->m);
      sync();
    } catch (...) {
      delete mm;
      throw;
    }
    return mm;
  });
}

int ohx_mesh_batch_adapt(ohx_mesh* m, double sf, double gq, unsigned nl, unsigned mo, int* results,
                         double* stats) {
  return guard(-2, [&]() {
    OHX_REQUIRE(m->batch, "batch: not a batch mesh");
    AdaptStats st;
    batch_adapt(m->m, *m->batch, sf, gq, nl, mo, results, &st);
    if (stats) {
      stats[0] = st.npass;
      stats[1] = st.nelems;
      stats[2] = st.tpass;
      stats[3] = st.ttotal;
    }
    int limited = 0;
    for (u32 c = 0; c < m->batch->ncomp; ++c) limited += results[c] < 0;
    return limited;
  });
}

// Worker threads that adapt independent meshes concurrently, each on a
// stream of its own: a small mesh's pass is a chain of short kernels and
// host decisions, so meshes on several streams keep the GPU busy while each
// one waits on its own synchronisations.  The workers live until the
// process ends.
namespace {
struct Pool {
  std::mutex mu;
  std::condition_variable work, done_cv;
  std::vector<std::thread> workers;
  std::function<void(u32)> fn;
  u32 n = 0, next = 0, done = 0;
  u32 active = 0;  // workers 0 .. active-1 take part in the current run
  unsigned long long gen = 0;
  int device = 0;
  void grow(u32 k) {
    while (workers.size() < k) {
      u32 id = (u32) workers.size();
      workers.emplace_back([this, id] { loop(id); });
      workers.back().detach();
    }
  }
  void loop(u32 id) {
    (void) hipSetDevice(device);
    unsigned long long seen = 0;
    std::unique_lock<std::mutex> l(mu);
    for (;;) {
      work.wait(l, [&] { return gen != seen && next < n && id < active; });
      while (next < n) {
        u32 i = next++;
        l.unlock();
        fn(i);
        l.lock();
        if (++done == n) done_cv.notify_all();
      }
      seen = gen;
    }
  }
  void run(u32 k, u32 count, std::function<void(u32)> f) {
    std::unique_lock<std::mutex> l(mu);
    grow(k);
    fn = std::move(f);
    n = count;
    next = 0;
    done = 0;
    active = k;
    ++gen;
    work.notify_all();
    done_cv.wait(l, [&] { return done == n; });
  }
};
Pool& pool() {
  static Pool* p = new Pool;  // never destroyed: its threads outlive static teardown
  return *p;
}
}  // namespace

int ohx_adapt_many(unsigned n, ohx_mesh* const* meshes, double sf, double gq, unsigned nl, unsigned mo,
                   unsigned nthreads, int* results, double* stats) {
  return guard(-2, [&]() {
    OHX_REQUIRE(meshes && results, "adapt_many: null arguments");
    sync();  // the meshes' arrays are complete before other streams read them
    std::vector<AdaptStats> st(n);
    std::vector<std::string> err(n);
    Pool& p = pool();
    p.device = rt().device;
    // largest meshes first (by element count): the run ends when its last
    // mesh does, so a large mesh started late would leave the other streams
    // idle behind it
    std::vector<u32> order(n);
    for (u32 i = 0; i < n; ++i) order[i] = i;
    std::vector<u32> size(n);
    for (u32 i = 0; i < n; ++i) size[i] = meshes[i] ? count(meshes[i]->m, meshes[i]->m.dim) : 0;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return size[a] > size[b]; });
    auto t0 = std::chrono::steady_clock::now();
    p.run(nthreads ? nthreads : 1, n, [&](u32 k) {
      u32 const i = order[k];
      try {
        results[i] = adapt(meshes[i]->m, sf, gq, nl, mo, &st[i]);
      } catch (std::exception const& e) {
        results[i] = -2;
        err[i] = e.what();
      }
    });
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int bad = 0;
    std::string first;
    for (u32 i = 0; i < n; ++i)
      if (results[i] == -2) {
        if (!bad) first = err[i];
        ++bad;
      }
    if (stats) {
      stats[0] = stats[1] = stats[2] = 0;
      for (auto& s : st) {
        stats[0] += s.npass;
        stats[1] += s.nelems;
        stats[2] += s.tpass;
      }
      stats[3] = wall;
    }
    if (bad) fail(("adapt_many: " + first).c_str(), __FILE__, __LINE__);
    int limited = 0;
    for (u32 i = 0; i < n; ++i) limited += results[i] < 0;
    return limited;
  });
}

ohx_mesh* ohx_mesh_read_vtk(char const* path) {
  return guard((ohx_mesh*) nullptr, [&]() {
    auto* mm = new ohx_mesh(3, false);
    try {
      read_vtk(path, mm->m);
      sync();
    } catch (...) {
      delete mm;
      throw;
    }
    return mm;
  });
}

int ohx_mesh_write_vtk(ohx_mesh* m, char const* path) {
  return guard(-1, [&]() {
    write_vtk(m->m, path);
    return 0;
  });
}

void ohx_set_vtk_compression(int on) { g_vtk_compress = on != 0; }

unsigned ohx_mesh_diff(ohx_mesh* a, ohx_mesh* b, double tol, double floor, unsigned allow_superset) {
  return guard(OHX_ERROR, [&]() {
    u32 r = mesh_diff(a->m, b->m, tol, floor, allow_superset != 0);
    std::fflush(stdout);
    return r;
  });
}

unsigned ohx_mesh_warp_to_limit(ohx_mesh* m, double qf) {
  return guard(OHX_ERROR, [&]() { return warp_to_limit(m->m, qf); });
}

osh_t ohx_mesh_osh(ohx_mesh* m) { return reinterpret_cast<osh_t>(m); }

void ohx_set_precision(int mixed) { g_mixed = mixed != 0; }
void ohx_mixed_stats(double* max_err, unsigned long long* screened, unsigned long long* reeval) {
  guard_void([&]() { mixed_stats(max_err, screened, reeval); });
}
void ohx_mixed_stats_reset(void) {
  guard_void([&]() { mixed_stats_reset(); });
}
void ohx_tuning_force(char const* name, int variant) {
  guard_void([&]() { tuning_force(name, variant); });
}
unsigned ohx_tuning_load(char const* path) {
  return guard(OHX_ERROR, [&]() { return tuning_load_file(path); });
}
unsigned ohx_tuning_report(char* buf, unsigned cap) {
  return guard(0u, [&]() {
    std::string r = tuning_report();
    if (buf && cap) {
      size_t n = r.size() < cap - 1 ? r.size() : cap - 1;
      std::memcpy(buf, r.data(), n);
      buf[n] = 0;
    }
    return (unsigned) r.size();
  });
}
void ohx_set_indset_max_rounds(unsigned n) { g_indset_max_rounds = n ? n : 100; }
unsigned ohx_indset_rounds_max(int reset) {
  return reset ? g_indset_rounds_seen.exchange(0) : g_indset_rounds_seen.load();
}
void ohx_set_verbose(int on) { g_verbose = on; }
void ohx_profile_enable(int on) { prof_enable(on != 0); }
void ohx_profile_only(char const* name) { prof_only(name); }
void ohx_profile_reset(void) {
  guard_void([&]() { prof_reset(); });
}
int ohx_profile_get(char const* name, double* ms, unsigned long long* launches, double* units) {
  return guard(0, [&]() { return prof_get(name, ms, launches, units) ? 1 : 0; });
}
unsigned ohx_profile_names(char* buf, unsigned cap) {
  return guard(0u, [&]() -> unsigned {
    std::string all;
    for (auto& n : prof_names()) all += n + "\n";
    if (buf && cap) {
      std::strncpy(buf, all.c_str(), cap - 1);
      buf[cap - 1] = 0;
    }
    return (unsigned) all.size();
  });
}
unsigned long long ohx_device_bytes_high_water(void) { return dev_bytes_high_water(); }
void ohx_device_reset_high_water(void) { dev_reset_high_water(); }
void ohx_synchronize(void) {
  guard_void([&]() { sync(); });
}
unsigned ohx_swap_nint_edges(unsigned rs) {
  return guard(OHX_ERROR, [&]() { return swap_nint_edges(rs); });
}
void ohx_swap_int_edges(unsigned rs, unsigned code, unsigned* out) {
  guard_void([&]() { swap_int_edges_host(rs, code, out); });
}

// ---------------------------------------------------------- public osh_*
void osh_init(int*, char***) {
  osh_guard([]() {
    rt();
    // the fixed scan scratch of ohx_exscan_u32_noalloc on this device
    noalloc_setup_current_device();
    return 0;
  });
}

void osh_fini(void) {
  try {
    dev_trim();
  } catch (...) {
  }
}

void osh_free(osh_t m) { delete M(m); }

osh_t osh_read_vtk(char const* filename) {
  return osh_guard([&]() {
    auto* mm = new ohx_mesh(3, false);
    try {
      read_vtk(filename, mm->m);
      sync();
    } catch (...) {
      delete mm;
      throw;
    }
    return reinterpret_cast<osh_t>(mm);
  });
}

void osh_write_vtk(osh_t m, char const* filename) {
  osh_guard([&]() {
    write_vtk(M(m)->m, filename);
    return 0;
  });
}

osh_t osh_new(unsigned dim) {
  return reinterpret_cast<osh_t>(new ohx_mesh(dim, true));
}

unsigned* osh_build_ents(osh_t m, unsigned d, unsigned n) {
  return osh_guard([&]() -> unsigned* {
    Mesh& mm = M(m)->m;
    if (d == 0) {
      set_ents(mm, 0, n, Buf<u32>());
      return nullptr;
    }
    Buf<u32> b((size_t) n * h_down_deg[d][0]);
    u32* p = b.get();
    set_ents(mm, d, n, std::move(b));
    return p;
  });
}

unsigned osh_dim(osh_t m) { return M(m)->m.dim; }
unsigned osh_nelems(osh_t m) { return osh_count(m, M(m)->m.dim); }
unsigned osh_nverts(osh_t m) { return osh_count(m, 0); }
unsigned osh_count(osh_t m, unsigned d) {
  return osh_guard([&]() { return count(M(m)->m, d); });
}

unsigned const* osh_down(osh_t m, unsigned h, unsigned l) {
  return osh_guard([&]() { return ask_down(M(m)->m, h, l); });
}
unsigned const* osh_up(osh_t m, unsigned l, unsigned h) {
  return osh_guard(
      [&]() -> unsigned const* { return ask_up(M(m)->m, l, h).adj.get(); });
}
unsigned const* osh_up_offs(osh_t m, unsigned l, unsigned h) {
  return osh_guard(
      [&]() -> unsigned const* { return ask_up(M(m)->m, l, h).offsets.get(); });
}
unsigned const* osh_up_dirs(osh_t m, unsigned l, unsigned h) {
  return osh_guard(
      [&]() -> unsigned const* { return ask_up(M(m)->m, l, h).dirs.get(); });
}
unsigned const* osh_star(osh_t m, unsigned l, unsigned h) {
  return osh_guard(
      [&]() -> unsigned const* { return ask_star(M(m)->m, l, h).adj.get(); });
}
unsigned const* osh_star_offs(osh_t m, unsigned l, unsigned h) {
  return osh_guard(
      [&]() -> unsigned const* { return ask_star(M(m)->m, l, h).offsets.get(); });
}
double const* osh_coords(osh_t m) {
  return osh_guard([&]() { return coords(M(m)->m); });
}

unsigned const* osh_own_rank(osh_t m, unsigned d) {
  return osh_guard([&]() -> unsigned const* {
    ohx_mesh* mm = M(m);
    mm->own_rank[d] = filled(count(mm->m, d), 0);
    return mm->own_rank[d].get();
  });
}
unsigned const* osh_own_id(osh_t m, unsigned d) {
  return osh_guard([&]() -> unsigned const* {
    ohx_mesh* mm = M(m);
    mm->own_id[d] = iota(count(mm->m, d));
    return mm->own_id[d].get();
  });
}
unsigned long const* osh_global(osh_t m, unsigned d) {
  return osh_guard([&]() -> unsigned long const* {
    ohx_mesh* mm = M(m);
    u32 n = count(mm->m, d);
    mm->global[d] = Buf<unsigned long>(n);
    unsigned long* g = mm->global[d].get();
    parallel_for(n, OHX_LAMBDA(u32 i) { g[i] = i; });
    return g;
  });
}

double* osh_new_field(osh_t m, unsigned d, char const* name, unsigned nc, enum osh_transfer tt) {
  return osh_guard([&]() -> double* {
    Mesh& mm = M(m)->m;
    if (Tag* t = find_tag(mm, d, name)) return t->f.get();
    return add_tag_f64(mm, d, name, nc, Buf<double>((size_t) count(mm, d) * nc), (int) tt).f.get();
  });
}
double* osh_get_field(osh_t m, unsigned d, char const* name) {
  return osh_guard([&]() -> double* {
    Tag* t = find_tag(M(m)->m, d, name);
    OHX_REQUIRE(t && t->type == TagType::F64, "no such field");
    return t->f.get();
  });
}
void osh_free_field(osh_t m, unsigned d, char const* name) {
  osh_guard([&]() {
    remove_tag(M(m)->m, d, name);
    return 0;
  });
}
unsigned osh_nfields(osh_t m, unsigned d) {
  unsigned n = 0;
  for (auto& t : M(m)->m.tags[d]) n += (t->type == TagType::F64);
  return n;
}
char const* osh_field(osh_t m, unsigned d, unsigned i) {
  unsigned n = 0;
  for (auto& t : M(m)->m.tags[d])
    if (t->type == TagType::F64 && i == n++) return t->name.c_str();
  return nullptr;
}
unsigned osh_components(osh_t m, unsigned d, char const* name) {
  Tag* t = find_tag(M(m)->m, d, name);
  return t ? t->ncomps : 0;
}
unsigned* osh_new_label(osh_t m, unsigned d, char const* name, unsigned nc) {
  return osh_guard([&]() -> unsigned* {
    Mesh& mm = M(m)->m;
    if (Tag* t = find_tag(mm, d, name)) return t->u.get();
    return add_tag_u32(mm, d, name, nc, Buf<u32>((size_t) count(mm, d) * nc)).u.get();
  });
}
unsigned* osh_get_label(osh_t m, unsigned d, char const* name) {
  return osh_guard([&]() -> unsigned* {
    Tag* t = find_tag(M(m)->m, d, name);
    OHX_REQUIRE(t && t->type == TagType::U32, "no such label");
    return t->u.get();
  });
}
void osh_free_label(osh_t m, unsigned d, char const* name) { osh_free_field(m, d, name); }

unsigned long* osh_new_global(osh_t m, unsigned d) {
  return const_cast<unsigned long*>(osh_global(m, d));
}
void osh_accumulate_to_owner(osh_t, unsigned, char const*) {}
void osh_conform(osh_t, unsigned, char const*) {}

void osh_mark_classified(osh_t m, unsigned ent_dim, unsigned class_dim, unsigned class_id, unsigned* marked) {
  osh_guard([&]() {
    Buf<u32> mk = mark_class(M(m)->m, ent_dim, class_dim, class_id);
    u32 const* k = mk.get();
    parallel_for(count(M(m)->m, ent_dim), OHX_LAMBDA(u32 i) {
      if (k[i]) marked[i] = 1;
    });
    sync();
    return 0;
  });
}

void osh_ghost(osh_t, unsigned nlayers) {
  if (nlayers) {
    std::fprintf(stderr, "omega_h (MI355X): ghost layers need MPI ranks; this build holds the mesh on one GPU\n");
    std::abort();
  }
}
unsigned osh_num_ghost_layers(osh_t) { return 0; }

unsigned osh_adapt(osh_t m, double sf, double gq, unsigned nl, unsigned mp) {
  return osh_guard([&]() -> unsigned {
    int r = adapt(M(m)->m, sf, gq, nl, mp, nullptr);
    if (r < 0) {
      std::fprintf(stderr, "mesh_adapt could not succeed after %u operations\n", mp);
      std::abort();
    }
    return (unsigned) r;
  });
}

void osh_identity_size(osh_t m, char const* name) {
  osh_guard([&]() {
    identity_size(M(m)->m, name);
    return 0;
  });
}

}  // extern "C"
