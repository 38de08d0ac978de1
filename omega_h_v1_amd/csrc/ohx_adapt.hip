// Refine, coarsen and swap passes, the independent set, the adapt driver
// and warp_to_limit, on a device-resident mesh.  Every per-entity loop is
// a gfx950 kernel; the host only sequences passes and reads counts.
#include <chrono>
#include <cstring>

#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_prims.hpp"
#include "ohx_tables.hpp"

namespace ohx {

int g_verbose = 0;

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// scratch word slots (zeroed by a fill kernel before each use)
enum { W_ANY = 0, W_ERR = 1, W_FLAGS = 8 };

static u32 read_word(u32 k) { return read_at(scratch_words(), k); }
static void zero_words(u32 first, u32 n) { fill(scratch_words() + first, n, 0); }

// ------------------------------------------------------- independent set
// One Jacobi round of the deterministic Luby variant (indset.cpp:55-84):
// an undecided entity leaves the set if a neighbour is in it, waits if a
// live neighbour has larger goodness or equal goodness and a lower index,
// and joins otherwise.  flag is set if anything is still undecided.
enum { NOT_IN_SET = 0, IN_SET = 1, UNKNOWN = 2 };

__global__ __launch_bounds__(BLOCK) void indset_round(u32 n, u32 const* __restrict__ offs,
                                                      u32 const* __restrict__ adj,
                                                      double const* __restrict__ good,
                                                      u32 const* __restrict__ old, u32* __restrict__ state,
                                                      u32* __restrict__ flag) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 s = old[i];
  if (s != UNKNOWN) {
    state[i] = s;
    return;
  }
  u32 f = offs[i], e = offs[i + 1];
  for (u32 j = f; j < e; ++j)
    if (old[adj[j]] == IN_SET) {
      state[i] = NOT_IN_SET;
      return;
    }
  double myg = good[i];
  for (u32 j = f; j < e; ++j) {
    u32 o = adj[j];
    if (old[o] == NOT_IN_SET) continue;
    double og = good[o];
    if ((myg == og && o < i) || myg < og) {
      state[i] = UNKNOWN;
      *flag = 1u;
      return;
    }
  }
  state[i] = IN_SET;
}

Buf<u32> find_indset(Mesh& m, u32 ent_dim, u32 const* filter, double const* good) {
  u32 n = count(m, ent_dim);
  Graph const& g = ask_star(m, ent_dim, m.dim);
  Buf<u32> a(n), b(n);
  {
    u32* s = a.get();
    parallel_for(n, OHX_LAMBDA(u32 i) { s[i] = filter[i] ? UNKNOWN : NOT_IN_SET; });
  }
  if (!n) return a;
  // per round: state in and out, goodness and the star CSR
  ProfScope ps("indset", (double) n * 16 + (double) g.adj.size() * 4);
  // rounds are launched in batches; a round after convergence copies the
  // state unchanged, so running past the converged round is harmless
  constexpr u32 BATCH = 4, MAX_ROUNDS = 100;
  u32* flags = scratch_words() + W_FLAGS;
  for (u32 done = 0; done < MAX_ROUNDS; done += BATCH) {
    fill(flags, BATCH, 0);
    for (u32 r = 0; r < BATCH; ++r) {
      hipLaunchKernelGGL(indset_round, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, g.offsets.get(),
                         g.adj.get(), good, a.get(), b.get(), flags + r);
      std::swap(a, b);
    }
    check_launch("indset_round");
    u32 h[BATCH];
    to_host(h, flags, BATCH);
    if (!h[BATCH - 1]) return a;
  }
  fail("independent set did not converge in 100 rounds (indset.cpp:107-116)", __FILE__, __LINE__);
  return a;
}

// ------------------------------------------------------------- inherit
// class tags of products come from their domains (inherit.cpp:14-232)
struct Prods {
  u32 ndoms[4] = {0, 0, 0, 0};
  u32 const* offs[4] = {nullptr, nullptr, nullptr, nullptr};
  Buf<u32> own[4];
  u32 ngen[4] = {0, 0, 0, 0};  // host copies of offs[d][ndoms[d]]
  u32 go[5] = {0, 0, 0, 0, 0};
  void finish() {
    go[0] = 0;
    for (u32 i = 0; i < 4; ++i) {
      ngen[i] = (ndoms[i] && offs[i]) ? read_at(offs[i], ndoms[i]) : 0;
      go[i + 1] = go[i] + ngen[i];
    }
  }
};

static void inherit_tag(Mesh& m, Mesh& mo, u32 prod_dim, Prods const& p, char const* name) {
  Tag* t = find_tag(m, prod_dim, name);
  u32 w = t->ncomps;
  Buf<u32> out((size_t) p.go[4] * w);
  for (u32 d = 0; d < 4; ++d) {
    u32 const* src = nullptr;
    if (d == 0)
      src = t->u.get();
    else if (d <= m.dim) {
      Tag* td = find_tag(m, d, name);
      src = td ? td->u.get() : nullptr;
    }
    if (!(p.ndoms[d] && src && p.offs[d])) continue;
    expand_into_u32(out.get() + (size_t) p.go[d] * w, src, p.offs[d], p.ndoms[d], w);
  }
  add_tag_u32(mo, prod_dim, name, w, std::move(out));
}

static void inherit_class(Mesh& m, Mesh& mo, u32 prod_dim, Prods const& p) {
  if (!find_tag(m, 0, "class_dim")) return;
  OHX_REQUIRE(m.full, "class inheritance needs a full mesh");
  inherit_tag(m, mo, prod_dim, p, "class_dim");
  if (find_tag(m, 0, "class_id")) inherit_tag(m, mo, prod_dim, p, "class_id");
}

// --------------------------------------------------------------- refine
// Worst child quality of every candidate edge split
// (refine_qualities.cpp:16-82); rejected candidates are cleared.
template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void refine_quality_f64(
    u32 ns, u32 const* __restrict__ vs, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
    u32 const* __restrict__ ud, u32 const* __restrict__ ve, double const* __restrict__ x,
    double const* __restrict__ eq, double qual_floor, u32* __restrict__ cand, double* __restrict__ out,
    u32* __restrict__ words) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= ns || !cand[i]) return;
  constexpr u32 VPE = DIM + 1;
  V3 a = ld3(x + 3 * (size_t) vs[2 * i]), b = ld3(x + 3 * (size_t) vs[2 * i + 1]);
  // average_element_field (element_field.hpp:8-25): ((0 + a) + b) / 2
  V3 split;
  {
    double s;
    s = 0; s += a.x; s += b.x; s /= 2; split.x = s;
    s = 0; s += a.y; s += b.y; s /= 2; split.y = s;
    s = 0; s += a.z; s += b.z; s /= 2; split.z = s;
  }
  double minq = 1, old_minq = 1;
  u32 const* bo = opp(DIM, 0);
  for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
    u32 dir = ud[j], e = ua[j];
    if (eq && eq[e] < old_minq) old_minq = eq[e];
    u32 const* evs = canon(DIM, 1, 0, dir);
    for (u32 k = 0; k < 2; ++k) {
      u32 const* bv = canon(DIM, DIM - 1, 0, bo[evs[k]]);
      V3 p[VPE];
#pragma unroll
      for (u32 l = 0; l + 1 < VPE; ++l) p[l] = ld3(x + 3 * (size_t) ve[(size_t) e * VPE + bv[l]]);
      p[VPE - 1] = split;
      double q = entity_quality(DIM, p);
      if (!(q > 0)) words[W_ERR] = 1u;
      if (q < minq) minq = q;
    }
  }
  if ((minq < qual_floor) || (eq && (minq <= old_minq)))
    cand[i] = 0;
  else {
    out[i] = minq;
    words[W_ANY] = 1u;
  }
}

// candidates of refine_by_size (refine.cpp:11-15), fused with measure_edge
__global__ __launch_bounds__(BLOCK) void refine_candidate_f64(u32 n, u32 const* __restrict__ ev,
                                                              double const* __restrict__ x,
                                                              double const* __restrict__ h,
                                                              u32* __restrict__ cand, u32* __restrict__ words) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 a = ev[2 * i], b = ev[2 * i + 1];
  double len = sqrt(sqdist(ld3(x + 3 * (size_t) b), ld3(x + 3 * (size_t) a)));
  double want = (h[a] + h[b]) / 2;
  u32 c = (len / want) > 1.0 ? 1u : 0u;
  cand[i] = c;
  if (c) words[W_ANY] = 1u;
}

static Buf<double> refine_qualities(Mesh& m, u32* cand, double qual_floor, bool require_better) {
  u32 dim = m.dim;
  u32 ns = count(m, 1);
  u32 const* vs = ask_down(m, 1, 0);
  u32 const* ve = ask_down(m, dim, 0);
  Up const& u = ask_up(m, 1, dim);
  Buf<double> eq;
  if (require_better) eq = mesh_qualities(m);
  Buf<double> out(ns);
  zero_words(0, 2);
  ProfScope ps("refine_quality_f64", (double) ns * 20 + (double) u.adj.size() * 8 + (double) count(m, dim) * (dim + 1) * 4 + (double) count(m, 0) * 24);
  if (dim == 3)
    hipLaunchKernelGGL(refine_quality_f64<3>, dim3(nblocks(ns)), dim3(BLOCK), 0, stream(), ns, vs,
                       u.offsets.get(), u.adj.get(), u.dirs.get(), ve, coords(m), eq.get(), qual_floor,
                       cand, out.get(), scratch_words());
  else
    hipLaunchKernelGGL(refine_quality_f64<2>, dim3(nblocks(ns)), dim3(BLOCK), 0, stream(), ns, vs,
                       u.offsets.get(), u.adj.get(), u.dirs.get(), ve, coords(m), eq.get(), qual_floor,
                       cand, out.get(), scratch_words());
  check_launch("refine_quality_f64");
  return out;
}

// refine_verts + refine_node (refine_common.cpp:29-54, refine_nodal.cpp:10-46)
static void refine_verts(Mesh& m, Mesh& mo, u32 const* gen_off, u32 nsplit) {
  u32 nv = count(m, 0), ns = count(m, 1);
  set_ents(mo, 0, nv + nsplit, Buf<u32>());
  u32 const* vs = ask_down(m, 1, 0);
  for (auto& tp : m.tags[0]) {
    Tag& t = *tp;
    if (t.type != TagType::F64) continue;
    u32 w = t.ncomps;
    Buf<double> out((size_t) (nv + nsplit) * w);
    d2d(out.get(), t.f.get(), (size_t) nv * w);
    double const* f = t.f.get();
    double* o = out.get();
    parallel_for(ns, OHX_LAMBDA(u32 i) {
      u32 g = gen_off[i];
      if (g == gen_off[i + 1]) return;
      double* dst = o + (size_t) (nv + g) * w;
      u32 v0 = vs[2 * i], v1 = vs[2 * i + 1];
      for (u32 c = 0; c < w; ++c) {
        double s = 0;
        s += f[(size_t) v0 * w + c];
        s += f[(size_t) v1 * w + c];
        s /= 2;
        dst[c] = s;
      }
    });
    add_tag_f64(mo, 0, t.name.c_str(), w, std::move(out), t.transfer);
  }
}

static u32 prods_per_dom(u32 dom_dim, u32 src_dim, u32 prod_dim) {  // refine_topology.cpp:110-121
  if (prod_dim == 0) return (dom_dim == src_dim) ? 1 : 0;
  u32 base_dim = prod_dim - 1;
  u32 opp_dim = dom_dim - 1 - base_dim;
  if (src_dim < opp_dim) return 0;
  return h_down_deg[src_dim][opp_dim];
}

// refine_ents (refine_common.cpp:64-129) with split_to_domain
// (splits_to_domains.cpp:13-56) and refine_domain_entity
// (refine_topology.cpp:26-107)
static void refine_ents(Mesh& m, Mesh& mo, u32 const* gen_off) {
  u32 dim = m.dim;
  u32 nv = count(m, 0);
  Buf<u32> off[4], dir[4], vert[4];
  off[0] = filled(nv + 1, 0);
  for (u32 d = 1; d <= dim; ++d) {
    if (!has_dim(m, d)) continue;
    u32 nd = count(m, d);
    u32 const* sd = ask_down(m, d, 1);
    u32 spd = h_down_deg[d][1];
    Buf<u32> will(nd);
    dir[d] = Buf<u32>(nd);
    vert[d] = Buf<u32>(nd);
    u32* wl = will.get();
    u32* dr = dir[d].get();
    u32* vt = vert[d].get();
    parallel_for(nd, OHX_LAMBDA(u32 i) {
      u32 w = 0, dd = 0, vv = 0;
      for (u32 j = 0; j < spd; ++j) {
        u32 s = sd[(size_t) i * spd + j];
        if (gen_off[s] == gen_off[s + 1]) continue;
        w = 1;
        dd = j;
        vv = nv + gen_off[s];
        break;
      }
      wl[i] = w;
      dr[i] = dd;
      vt[i] = vv;
    });
    off[d] = exscan(wl, nd);
  }
  for (u32 pd = 0; pd <= dim; ++pd) {
    if (!m.full && 0 < pd && pd < dim) continue;
    // setup_refine (inherit.cpp:104-138)
    Prods p;
    u32 np = count(m, pd);
    p.ndoms[0] = np;
    p.own[0] = negate_offsets(off[pd].get(), np);
    p.offs[0] = p.own[0].get();
    u32 lo, hi;
    if (m.full) {
      lo = pd > 1 ? pd : 1;
      hi = dim;
    } else {
      lo = hi = (pd == 0) ? 1 : dim;
    }
    for (u32 d = lo; d <= hi; ++d) {
      p.ndoms[d] = count(m, d);
      p.own[d] = scale(off[d].get(), p.ndoms[d] + 1, prods_per_dom(d, 1, pd));
      p.offs[d] = p.own[d].get();
    }
    p.finish();
    if (pd) {
      u32 nout = p.go[4];
      u32 vpp = pd + 1;
      Buf<u32> vo((size_t) nout * vpp);
      expand_into_u32(vo.get(), ask_down(m, pd, 0), p.offs[0], np, vpp);
      for (u32 d = 1; d <= dim; ++d) {
        if (!(p.ndoms[d] && p.offs[d]) || !p.ngen[d]) continue;
        u32 base_dim = pd - 1;
        u32 opp_dim = d - 1 - base_dim;
        if (1 < opp_dim) continue;
        u32 ops = h_down_deg[1][opp_dim];
        u32 vpd = d + 1, vpb = vpp - 1;
        u32 const* vd = ask_down(m, d, 0);
        u32 const* of = off[d].get();
        u32 const* dr = dir[d].get();
        u32 const* vt = vert[d].get();
        u32* out = vo.get() + (size_t) p.go[d] * vpp;
        parallel_for(count(m, d), OHX_LAMBDA(u32 i) {
          u32 o = of[i];
          if (o == of[i + 1]) return;
          u32 const* os = canon(d, 1, opp_dim, dr[i]);
          u32 const* bo = opp(d, opp_dim);
          u32* q = out + (size_t) o * ops * vpp;
          for (u32 j = 0; j < ops; ++j) {
            u32 const* bv = canon(d, base_dim, 0, bo[os[j]]);
            for (u32 k = 0; k < vpb; ++k) q[k] = vd[(size_t) i * vpd + bv[k]];
            q[vpb] = vt[i];
            q += vpp;
          }
        });
      }
      set_ents(mo, pd, nout, std::move(vo));
    }
    inherit_class(m, mo, pd, p);
  }
}

static void say(char const* verb, u32 n, u32 d) {
  static char const* const names[4][2] = {{"vertex", "vertices"}, {"edge", "edges"},
                                          {"triangle", "triangles"}, {"tet", "tets"}};
  if (g_verbose) std::printf("%s %10u %s\n", verb, n, names[d][n != 1]);
}

// refine_common (refine_common.cpp:131-190) for edges
static u32 refine_common(Mesh& m, Buf<u32>& cand, double qual_floor, bool require_better) {
  u32 ns = count(m, 1);
  Buf<double> q = refine_qualities(m, cand.get(), qual_floor, require_better);
  u32 words[2];
  to_host(words, scratch_words(), 2);
  OHX_REQUIRE(!words[W_ERR], "refine: a child element has non-positive quality (refine_qualities.cpp:72)");
  if (!words[W_ANY]) return 0;
  Buf<u32> indset = find_indset(m, 1, cand.get(), q.get());
  cand.release();
  q.release();
  Buf<u32> gen_off = exscan(indset.get(), ns);
  indset.release();
  u32 total = read_at(gen_off.get(), ns);
  Mesh mo(m.dim, m.full);
  refine_verts(m, mo, gen_off.get(), total);
  refine_ents(m, mo, gen_off.get());
  say("split", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 refine_by_size(Mesh& m, double qual_floor) {
  u32 ne = count(m, 1);
  Tag* st = find_tag(m, 0, "adapt_size");
  OHX_REQUIRE(st && st->type == TagType::F64, "mesh has no adapt_size field");
  Buf<u32> cand(ne);
  zero_words(0, 1);
  if (ne) {
    ProfScope ps("refine_candidate_f64", (double) ne * 12 + (double) count(m, 0) * 32);
    hipLaunchKernelGGL(refine_candidate_f64, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne,
                       ask_down(m, 1, 0), coords(m), st->f.get(), cand.get(), scratch_words());
    check_launch("refine_candidate_f64");
  }
  if (!read_word(W_ANY)) return 0;
  return refine_common(m, cand, qual_floor, false);
}

void uniformly_refine(Mesh& m) {
  Buf<u32> cand = filled(count(m, 1), 1);
  if (!count(m, 1)) return;
  refine_common(m, cand, 0.0, false);
}

// -------------------------------------------------------------- coarsen
enum { DONT_COLLAPSE = 0, COLLAPSE_BOTH = 3 };  // collapse_codes.hpp:8-31
__host__ __device__ inline u32 collapses(u32 c, u32 d) { return (c & (1u << d)) != 0; }
__host__ __device__ inline u32 dont_collapse(u32 c, u32 d) { return c & ~(1u << d); }

// check_basic_edge_class (check_collapse_class.cpp:12-55)
static void check_basic_class(Mesh& m, u32* codes) {
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vcd = u32_tag(m, 0, "class_dim");
  u32 const* ecd = u32_tag(m, 1, "class_dim");
  parallel_for(count(m, 1), OHX_LAMBDA(u32 i) {
    u32 c = codes[i];
    if (c == DONT_COLLAPSE) return;
    u32 c0 = vcd[ev[2 * i]], c1 = vcd[ev[2 * i + 1]], ce = ecd[i];
    if (c0 == c1) {
      if (c0 != ce) codes[i] = DONT_COLLAPSE;
      return;
    }
    if (collapses(c, 0) && c0 != ce) c = dont_collapse(c, 0);
    if (collapses(c, 1) && c1 != ce) c = dont_collapse(c, 1);
    codes[i] = c;
  });
}

// check_for_exposed_side (check_collapse_class.cpp:57-140)
static void check_exposed_sides(Mesh& m, u32 ed, u32* codes) {
  Up const& u = ask_up(m, 1, ed);
  u32 const* vent = ask_down(m, ed, 0);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* sent = ask_down(m, ed, ed - 1);
  u32 const* ecd = u32_tag(m, ed, "class_dim");
  u32 const* scd = u32_tag(m, ed - 1, "class_dim");
  u32 vpe = ed + 1, spe = h_down_deg[ed][ed - 1];
  u32 const* uo = u.offsets.get();
  u32 const* ua = u.adj.get();
  u32 const* ud = u.dirs.get();
  u32* err = scratch_words() + W_ERR;
  fill(err, 1, 0);
  parallel_for(count(m, 1), OHX_LAMBDA(u32 i) {
    u32 c = codes[i];
    if (c == DONT_COLLAPSE) return;
    u32 const* so = opp(ed, 0);
    for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
      u32 ent = ua[j];
      u32 const* eve = canon(ed, 1, 0, ud[j]);
      for (u32 k = 0; k < 2; ++k) {
        if (!collapses(c, k)) continue;
        u32 va = ev[2 * i + (1 - k)];
        u32 l;
        for (l = 0; l < 2; ++l) {
          u32 vb = vent[(size_t) ent * vpe + eve[l]];
          if (va == vb) {
            u32 side = sent[(size_t) ent * spe + so[eve[l]]];
            if (scd[side] != ecd[ent]) c = dont_collapse(c, k);
            break;
          }
        }
        if (l == 2) *err = 1u;
      }
    }
    codes[i] = c;
  });
  OHX_REQUIRE(read_word(W_ERR) == 0, "check_for_exposed_side: edge vertex not in entity");
}

// coarsen_quality (coarsen_qualities.cpp:12-79)
template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void coarsen_quality_f64(
    u32 ne, u32 const* __restrict__ ev, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
    u32 const* __restrict__ ud, u32 const* __restrict__ vel, double const* __restrict__ x,
    double const* __restrict__ eq, double qual_floor, u32* __restrict__ codes, double* __restrict__ out,
    u32* __restrict__ words) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= ne) return;
  u32 c = codes[i];
  if (c == DONT_COLLAPSE) return;
  constexpr u32 VPE = DIM + 1;
  u32 const* bo = opp(DIM, 0);
  for (u32 j = 0; j < 2; ++j) {
    if (!collapses(c, j)) continue;
    double minq = 1, old_minq = 1;
    u32 col = ev[2 * i + j], gen = ev[2 * i + (1 - j)];
    V3 p[VPE];
    p[VPE - 1] = ld3(x + 3 * (size_t) gen);
    for (u32 k = uo[col]; k < uo[col + 1]; ++k) {
      u32 e = ua[k];
      u32 const* v = vel + (size_t) e * VPE;
      if (eq && eq[e] < old_minq) old_minq = eq[e];
      bool adj = false;
#pragma unroll
      for (u32 l = 0; l < VPE; ++l)
        if (v[l] == gen) adj = true;
      if (adj) continue;
      u32 const* bv = canon(DIM, DIM - 1, 0, bo[ud[k]]);
#pragma unroll
      for (u32 l = 0; l + 1 < VPE; ++l) p[l] = ld3(x + 3 * (size_t) v[bv[l]]);
      double q = element_quality(DIM, p);
      if (q < minq) minq = q;
    }
    if ((minq < qual_floor) || (eq && (minq <= old_minq)))
      c = dont_collapse(c, j);
    else
      out[2 * i + j] = minq;
  }
  codes[i] = c;
  if (c != DONT_COLLAPSE) words[W_ANY] = 1u;
}

// coarsen_size_code (coarsen.cpp:13-21), fused with measure_edge
__global__ __launch_bounds__(BLOCK) void coarsen_size_code_f64(u32 n, u32 const* __restrict__ ev,
                                                               double const* __restrict__ x,
                                                               double const* __restrict__ h, double floor_,
                                                               u32* __restrict__ codes, u32* __restrict__ words) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 a = ev[2 * i], b = ev[2 * i + 1];
  double len = sqrt(sqdist(ld3(x + 3 * (size_t) b), ld3(x + 3 * (size_t) a)));
  double want = (h[a] + h[b]) / 2;
  u32 c = ((len / want) < floor_) ? (u32) COLLAPSE_BOTH : (u32) DONT_COLLAPSE;
  codes[i] = c;
  if (c) words[W_ANY] = 1u;
}

static u32 coarsen_common(Mesh& m, Buf<u32>& codes, double qual_floor, bool require_better) {
  u32 dim = m.dim;
  u32 ne = count(m, 1);
  check_basic_class(m, codes.get());
  for (u32 d = 2; d <= dim; ++d) check_exposed_sides(m, d, codes.get());
  if (max_u32(codes.get(), ne) == DONT_COLLAPSE) return 0;
  Buf<double> eq;
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vel = ask_down(m, dim, 0);
  Up const& ue = ask_up(m, 0, dim);
  if (require_better) eq = element_qualities(dim, count(m, dim), vel, coords(m));
  Buf<double> cq((size_t) 2 * ne);
  zero_words(0, 1);
  {
    ProfScope ps("coarsen_quality_f64", (double) ne * 32 + (double) ue.adj.size() * 8 + (double) count(m, dim) * (dim + 1) * 4 + (double) count(m, 0) * 28);
    if (dim == 3)
      hipLaunchKernelGGL(coarsen_quality_f64<3>, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, ev,
                         ue.offsets.get(), ue.adj.get(), ue.dirs.get(), vel, coords(m), eq.get(), qual_floor,
                         codes.get(), cq.get(), scratch_words());
    else
      hipLaunchKernelGGL(coarsen_quality_f64<2>, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, ev,
                         ue.offsets.get(), ue.adj.get(), ue.dirs.get(), vel, coords(m), eq.get(), qual_floor,
                         codes.get(), cq.get(), scratch_words());
    check_launch("coarsen_quality_f64");
  }
  eq.release();
  if (!read_word(W_ANY)) return 0;
  // collapse_to_vert (collapses_to_verts.cpp:13-70)
  u32 nv = count(m, 0);
  Up const& uv = ask_up(m, 0, 1);
  Buf<u32> cand(nv);
  Buf<double> vq(nv);
  {
    u32 const* uo = uv.offsets.get();
    u32 const* ua = uv.adj.get();
    u32 const* ud = uv.dirs.get();
    u32 const* cd = codes.get();
    double const* q = cq.get();
    u32* ca = cand.get();
    double* vqp = vq.get();
    parallel_for(nv, OHX_LAMBDA(u32 i) {
      double maxq = 0;
      u32 is = 0;
      for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
        u32 e = ua[j], d = ud[j];
        if (!collapses(cd[e], d)) continue;
        double qq = q[2 * (size_t) e + d];
        if (!is) {
          is = 1;
          maxq = qq;
        } else if (qq > maxq) {
          maxq = qq;
        }
      }
      ca[i] = is;
      vqp[i] = is ? maxq : 0.0;
    });
  }
  Buf<u32> indset = find_indset(m, 0, cand.get(), vq.get());
  cand.release();
  // vertex_dest (collapses_to_verts.cpp:72-135)
  Buf<u32> gvv(nv);
  {
    u32 const* uo = uv.offsets.get();
    u32 const* ua = uv.adj.get();
    u32 const* ud = uv.dirs.get();
    u32 const* cd = codes.get();
    double const* q = cq.get();
    double const* vqp = vq.get();
    u32 const* is = indset.get();
    u32* g = gvv.get();
    u32* err = scratch_words() + W_ERR;
    fill(err, 1, 0);
    parallel_for(nv, OHX_LAMBDA(u32 i) {
      if (!is[i]) {
        g[i] = INVALID;
        return;
      }
      double maxq = vqp[i];
      if (maxq == 1e10) maxq = -1e10;
      u32 gv = INVALID;
      for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
        u32 e = ua[j], d = ud[j];
        if (!collapses(cd[e], d)) continue;
        double qq = q[2 * (size_t) e + d];
        if (qq == 1e10) qq = -1e10;
        if (qq == maxq) {
          gv = ev[2 * (size_t) e + (1 - d)];
          break;
        }
      }
      if (gv == INVALID) *err = 1u;
      g[i] = gv;
    });
    OHX_REQUIRE(read_word(W_ERR) == 0, "vertex_dest: no destination (collapses_to_verts.cpp:100)");
  }
  codes.release();
  cq.release();
  vq.release();
  // coarsen_interior (coarsen_common.cpp:191-219)
  Buf<u32> goff = exscan(indset.get(), nv);
  u32 total = read_at(goff.get(), nv);
  indset.release();
  Buf<u32> soff = negate_offsets(goff.get(), nv);
  u32 nv_out = read_at(soff.get(), nv);
  Mesh mo(dim, m.full);
  set_ents(mo, 0, nv_out, Buf<u32>());
  for (auto& tp : m.tags[0]) {  // tags_subset (subset.cpp:15-43)
    Tag& t = *tp;
    if (t.type == TagType::F64) {
      Buf<double> o((size_t) nv_out * t.ncomps);
      expand_into_f64(o.get(), t.f.get(), soff.get(), nv, t.ncomps);
      add_tag_f64(mo, 0, t.name.c_str(), t.ncomps, std::move(o), t.transfer);
    } else {
      Buf<u32> o((size_t) nv_out * t.ncomps);
      expand_into_u32(o.get(), t.u.get(), soff.get(), nv, t.ncomps);
      add_tag_u32(mo, 0, t.name.c_str(), t.ncomps, std::move(o), t.transfer);
    }
  }
  // coarsen_all_ents (coarsen_common.cpp:91-107), top dimension first
  Buf<u32> fused[4];
  for (u32 dd = 0; dd < dim; ++dd) {
    u32 d = dim - dd;
    u32 n = count(m, d);
    u32 const* vent = ask_down(m, d, 0);
    u32 const* sent = ask_down(m, d, d - 1);
    u32 nsd = count(m, d - 1);
    fused[d - 1] = filled(nsd, 0);
    u32 vpe = d + 1, spe = h_down_deg[d][d - 1];
    Buf<u32> will(n), same(n), gv(n), gd(n);
    {
      // collapse_to_ent (collapses_to_ents.cpp:11-82)
      u32 const* go = goff.get();
      u32 const* g = gvv.get();
      u32 const* fe = fused[d] ? fused[d].get() : nullptr;
      u32* fs = fused[d - 1].get();
      u32* wl = will.get();
      u32* sm = same.get();
      u32* gvp = gv.get();
      u32* gdp = gd.get();
      parallel_for(n, OHX_LAMBDA(u32 i) {
        u32 const* v = vent + (size_t) i * vpe;
        u32 col = INVALID, dir = INVALID;
        for (u32 j = 0; j < vpe; ++j)
          if (go[v[j]] != go[v[j] + 1]) {
            col = v[j];
            dir = j;
          }
        wl[i] = 0;
        gvp[i] = 0;
        gdp[i] = 0;
        if (dir == INVALID) {
          sm[i] = 1;
          return;
        }
        sm[i] = 0;
        u32 target = g[col];
        u32 gdir = INVALID;
        for (u32 j = 0; j < vpe; ++j)
          if (v[j] == target) gdir = j;
        if (gdir != INVALID) {
          fs[sent[(size_t) i * spe + opp(d, 0)[gdir]]] = 1;
          return;
        }
        if (!(fe && fe[i])) {
          wl[i] = 1;
          gvp[i] = target;
          gdp[i] = dir;
        }
      });
    }
    Buf<u32> egoff = exscan(will.get(), n);
    Buf<u32> esoff = exscan(same.get(), n);
    will.release();
    same.release();
    u32 ngen = read_at(egoff.get(), n), nsame = read_at(esoff.get(), n);
    Buf<u32> vout((size_t) (nsame + ngen) * vpe);
    expand_into_u32(vout.get(), vent, esoff.get(), n, vpe);
    {
      // coarsen_elem (coarsen_topology.cpp:9-44): base opposite the
      // collapsing vertex, then the target vertex
      u32 const* eg = egoff.get();
      u32 const* gvp = gv.get();
      u32 const* gdp = gd.get();
      u32* out = vout.get() + (size_t) nsame * vpe;
      parallel_for(n, OHX_LAMBDA(u32 i) {
        u32 o = eg[i];
        if (o == eg[i + 1]) return;
        u32 const* bv = canon(d, d - 1, 0, opp(d, 0)[gdp[i]]);
        for (u32 j = 0; j + 1 < vpe; ++j) out[(size_t) o * vpe + j] = vent[(size_t) i * vpe + bv[j]];
        out[(size_t) o * vpe + vpe - 1] = gvp[i];
      });
    }
    {
      // remap_conn (coarsen_common.cpp:29-33)
      u32* vo = vout.get();
      u32 const* s = soff.get();
      parallel_for((nsame + ngen) * vpe, OHX_LAMBDA(u32 k) { vo[k] = s[vo[k]]; });
    }
    set_ents(mo, d, nsame + ngen, std::move(vout));
    Prods p;  // setup_coarsen (inherit.cpp:140-156)
    p.ndoms[0] = n;
    p.offs[0] = esoff.get();
    p.ndoms[d] = n;
    p.offs[d] = egoff.get();
    p.finish();
    inherit_class(m, mo, d, p);
  }
  say("collapsed", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 coarsen_by_size(Mesh& m, double qual_floor, double size_floor) {
  u32 ne = count(m, 1);
  Tag* st = find_tag(m, 0, "adapt_size");
  OHX_REQUIRE(st && st->type == TagType::F64, "mesh has no adapt_size field");
  Buf<u32> codes(ne);
  zero_words(0, 1);
  if (ne) {
    ProfScope ps("coarsen_size_code_f64", (double) ne * 12 + (double) count(m, 0) * 32);
    hipLaunchKernelGGL(coarsen_size_code_f64, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne,
                       ask_down(m, 1, 0), coords(m), st->f.get(), size_floor, codes.get(), scratch_words());
    check_launch("coarsen_size_code_f64");
  }
  if (!read_word(W_ANY)) return 0;
  return coarsen_common(m, codes, qual_floor, false);
}

// coarsen_slivers (coarsen.cpp:38-70)
u32 coarsen_slivers(Mesh& m, double qual_floor, u32 nlayers) {
  u32 dim = m.dim;
  Buf<u32> sl = mark_slivers(m, qual_floor, nlayers);
  Buf<u32> mv = mark_down(m, dim, 0, sl.get());
  sl.release();
  u32 ne = count(m, 1);
  u32 const* ev = ask_down(m, 1, 0);
  Buf<u32> codes(ne);
  u32* c = codes.get();
  u32 const* mk = mv.get();
  parallel_for(ne, OHX_LAMBDA(u32 i) {
    u32 r = 0;
    if (mk[ev[2 * i]]) r |= 1u;
    if (mk[ev[2 * i + 1]]) r |= 2u;
    c[i] = r;
  });
  if (max_u32(c, ne) == DONT_COLLAPSE) return 0;
  return coarsen_common(m, codes, 0.0, true);
}

// ----------------------------------------------------------------- swap
// Triangulations of an N-gon (N = 3..7), their unique triangles
// (edge_swap.cpp:21-200) and, generated on first use, the interior edges of
// each triangulation.  The triangulation index is the swap code; ties in
// quality keep the lowest code, so the order must equal the reference's.
static u32 const h_sw_sizes[8] = {0, 0, 0, 1, 2, 3, 4, 5};
static u32 const h_sw_counts[8] = {0, 0, 0, 1, 2, 5, 14, 42};
static u32 const h_sw_tri_off[8] = {0, 0, 0, 0, 1, 5, 15, 35};
static u32 const h_sw_mesh_off[8] = {0, 0, 0, 0, 1, 5, 20, 76};
static unsigned char const h_sw_tris[70][3] = {
    {0, 1, 2},  // N = 3
    {0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {1, 2, 3},  // N = 4
    {0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 1, 4}, {1, 3, 4}, {1, 2, 3}, {2, 3, 4}, {0, 2, 4}, {0, 1, 3},
    {1, 2, 4},  // N = 5
    {0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5}, {0, 2, 5}, {2, 4, 5}, {2, 3, 4}, {0, 3, 5}, {3, 4, 5},
    {0, 2, 4}, {2, 3, 5}, {1, 2, 3}, {0, 1, 3}, {0, 1, 5}, {1, 4, 5}, {1, 3, 4}, {0, 1, 4}, {1, 3, 5},
    {1, 2, 4}, {1, 2, 5},  // N = 6
    {0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5}, {0, 5, 6}, {0, 3, 6}, {3, 5, 6}, {3, 4, 5}, {0, 4, 6},
    {4, 5, 6}, {0, 3, 5}, {3, 4, 6}, {0, 2, 4}, {2, 3, 4}, {0, 2, 6}, {2, 5, 6}, {2, 4, 5}, {0, 2, 5},
    {2, 4, 6}, {2, 3, 5}, {2, 3, 6}, {0, 1, 3}, {1, 2, 3}, {0, 1, 4}, {1, 3, 4}, {0, 1, 6}, {1, 5, 6},
    {1, 4, 5}, {0, 1, 5}, {1, 4, 6}, {1, 3, 5}, {1, 3, 6}, {1, 2, 4}, {1, 2, 5}, {1, 2, 6}};  // N = 7
static unsigned char const h_sw_meshes[286] = {
    0,                                                                             // N = 3
    0, 1, 2, 3,                                                                    // N = 4
    0, 1, 2, 3, 4, 5, 0, 6, 7, 2, 5, 8, 3, 6, 9,                                   // N = 5
    0, 1, 2, 3, 0, 4, 5, 6, 0, 1, 7, 8, 0, 3, 6, 9, 0, 4, 8, 10, 2, 3, 11, 12, 11, 13, 14, 15, 7, 8, 11, 12,
    3, 11, 15, 16, 8, 11, 13, 17, 6, 13, 14, 18, 3, 6, 16, 18, 5, 6, 13, 19, 8, 10, 13, 19,  // N = 6
    0, 1, 2, 3, 4, 0, 1, 5, 6, 7, 0, 1, 2, 8, 9, 0, 1, 4, 7, 10, 0, 1, 5, 9, 11, 0, 3, 4, 12, 13, 0, 13, 14,
    15, 16, 0, 8, 9, 12, 13, 0, 4, 13, 16, 17, 0, 9, 13, 14, 18, 0, 7, 14, 15, 19, 0, 4, 7, 17, 19, 0, 6, 7,
    14, 20, 0, 9, 11, 14, 20, 2, 3, 4, 21, 22, 5, 6, 7, 21, 22, 2, 8, 9, 21, 22, 4, 7, 10, 21, 22, 5, 9, 11,
    21, 22, 3, 4, 22, 23, 24, 22, 24, 25, 26, 27, 8, 9, 22, 23, 24, 4, 22, 24, 27, 28, 9, 22, 24, 25, 29, 7,
    22, 25, 26, 30, 4, 7, 22, 28, 30, 6, 7, 22, 25, 31, 9, 11, 22, 25, 31, 3, 4, 13, 23, 32, 13, 25, 26, 27,
    32, 8, 9, 13, 23, 32, 4, 13, 27, 28, 32, 9, 13, 25, 29, 32, 13, 16, 25, 26, 33, 4, 13, 16, 28, 33, 13,
    15, 16, 25, 34, 9, 13, 18, 25, 34, 7, 19, 25, 26, 33, 4, 7, 19, 28, 33, 7, 15, 19, 25, 34, 6, 7, 20, 25,
    34, 9, 11, 20, 25, 34};  // N = 7

static __constant__ u32 k_sw_sizes[8] = {0, 0, 0, 1, 2, 3, 4, 5};
static __constant__ u32 k_sw_counts[8] = {0, 0, 0, 1, 2, 5, 14, 42};
static __constant__ u32 k_sw_tri_off[8] = {0, 0, 0, 0, 1, 5, 15, 35};
static __constant__ u32 k_sw_mesh_off[8] = {0, 0, 0, 0, 1, 5, 20, 76};
static __constant__ unsigned char k_sw_tris[70][3];
static __constant__ unsigned char k_sw_meshes[286];
// interior edges: offsets per ring size, then [code][edge][2]
static __constant__ u32 k_sw_nint[8];
static __constant__ u32 k_sw_int_off[8];
static __constant__ unsigned char k_sw_int[444];

static u32 h_sw_nint[8];
static u32 h_sw_int_off[8];
static unsigned char h_sw_int[444];
static bool g_swap_ready = false;

// The interior edges of each triangulation, derived the way the
// reference's generator does (test_print_swap_edges.cpp:29-64): the
// triangulation as a reduced 2D mesh, its derived edges, the non-boundary
// ones in edge order.  All triangulations of one N form one mesh of
// disjoint components; side derivation numbers each component's edges
// contiguously and in the same order as a mesh of its own would.
static void make_swap_tables() {
  if (g_swap_ready) return;
  u32 off = 0;
  for (u32 n = 0; n < 4; ++n) {
    h_sw_nint[n] = 0;
    h_sw_int_off[n] = 0;
  }
  for (u32 n = 4; n <= MAX_EDGE_SWAP; ++n) {
    u32 nm = h_sw_counts[n], nt = h_sw_sizes[n];
    std::vector<u32> conn((size_t) nm * nt * 3);
    for (u32 c = 0; c < nm; ++c)
      for (u32 k = 0; k < nt; ++k)
        for (u32 l = 0; l < 3; ++l)
          conn[((size_t) c * nt + k) * 3 + l] = c * n + h_sw_tris[h_sw_tri_off[n] + h_sw_meshes[h_sw_mesh_off[n] + c * nt + k]][l];
    Mesh m(2, false);
    set_ents(m, 0, nm * n, Buf<u32>());
    set_ents(m, 2, nm * nt, to_device(conn.data(), conn.size()));
    Buf<u32> bd = mark_part_boundary(m);
    u32 ne = count(m, 1);
    std::vector<u32> hb = host_copy(bd.get(), ne);
    std::vector<u32> ev = host_copy(ask_down(m, 1, 0), (size_t) 2 * ne);
    u32 nint_total = 0;
    for (u32 k = 0; k < ne; ++k) nint_total += !hb[k];
    OHX_REQUIRE(nint_total % nm == 0, "swap tables: uneven interior edge counts");
    u32 nint = nint_total / nm;
    h_sw_nint[n] = nint;
    h_sw_int_off[n] = off;
    std::vector<u32> seen(nm, 0);
    for (u32 k = 0; k < ne; ++k) {
      if (hb[k]) continue;
      u32 c = ev[2 * k] / n;
      OHX_REQUIRE(ev[2 * k + 1] / n == c && seen[c] < nint, "swap tables: bad interior edge");
      u32 slot = off + (c * nint + seen[c]) * 2;
      h_sw_int[slot] = (unsigned char) (ev[2 * k] - c * n);
      h_sw_int[slot + 1] = (unsigned char) (ev[2 * k + 1] - c * n);
      ++seen[c];
    }
    off += nm * nint * 2;
  }
  OHX_REQUIRE(off == 444, "swap tables: unexpected size");
  OHX_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(k_sw_tris), h_sw_tris, sizeof(h_sw_tris)));
  OHX_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(k_sw_meshes), h_sw_meshes, sizeof(h_sw_meshes)));
  OHX_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(k_sw_nint), h_sw_nint, sizeof(h_sw_nint)));
  OHX_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(k_sw_int_off), h_sw_int_off, sizeof(h_sw_int_off)));
  OHX_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(k_sw_int), h_sw_int, sizeof(h_sw_int)));
  g_swap_ready = true;
}

u32 swap_nint_edges(u32 rs) {
  make_swap_tables();
  return h_sw_nint[rs];
}

void swap_int_edges_host(u32 rs, u32 code, u32* out) {
  make_swap_tables();
  u32 k = h_sw_nint[rs];
  for (u32 i = 0; i < 2 * k; ++i) out[i] = h_sw_int[h_sw_int_off[rs] + code * k * 2 + i];
}

// find_edge_ring (edge_ring.cpp:12-74): ring vertices ordered by chaining
// each tet's opposite edge (a -> b), insertion style, <= 7 entries
__device__ inline u32 edge_ring(u32 edge, u32 const* uo, u32 const* ua, u32 const* ud, u32 const* ev,
                                u32 const* vt, u32* edge_v, u32* ring_v) {
  u32 f = uo[edge], e = uo[edge + 1];
  u32 rs = e - f;
  if (rs > MAX_EDGE_SWAP || rs < 3) return rs;
  u32 const* oe = opp(3, 1);
  edge_v[0] = ev[2 * (size_t) edge];
  edge_v[1] = ev[2 * (size_t) edge + 1];
  u32 ra[MAX_EDGE_SWAP], rb[MAX_EDGE_SWAP];
  for (u32 i = f; i < e; ++i) {
    u32 const* tv = vt + 4 * (size_t) ua[i];
    u32 te = ud[i];
    u32 const* tve = canon(3, 1, 0, te);
    u32 al = (tv[tve[0]] != edge_v[0]) ? 1u : 0u;
    u32 const* tvo = canon(3, 1, 0, oe[te]);
    ra[i - f] = tv[tvo[al]];
    rb[i - f] = tv[tvo[al ^ 1u]];
  }
  for (u32 i = 0; i + 1 < rs; ++i) {
    u32 j;
    for (j = i + 1; j < rs; ++j)
      if (ra[j] == rb[i]) break;
    if (j == rs) return 0;  // ring does not close: reported by the caller
    u32 ta = ra[j], tb = rb[j];
    ra[j] = ra[i + 1];
    rb[j] = rb[i + 1];
    ra[i + 1] = ta;
    rb[i + 1] = tb;
  }
  for (u32 i = 0; i < rs; ++i) ring_v[i] = ra[i];
  return rs;
}

// choose_edge_swap (edge_swap.cpp:230-280)
__device__ inline u32 choose_edge_swap(u32 rs, V3 const* ex, V3 const* rx, double* quality) {
  u32 tpm = k_sw_sizes[rs], nm = k_sw_counts[rs];
  unsigned char const* mesh = k_sw_meshes + k_sw_mesh_off[rs];
  unsigned char const(*tris)[3] = k_sw_tris + k_sw_tri_off[rs];
  u64 cached = 0;
  double cache[35];
  u32 code = INVALID;
  double best = -1;
  for (u32 i = 0; i < nm; ++i) {
    double mesh_minq = 1;
    for (u32 j = 0; j < tpm; ++j) {
      u32 tri = mesh[j];
      if (!((cached >> tri) & 1ull)) {
        V3 t0 = rx[tris[tri][0]], t1 = rx[tris[tri][1]], t2 = rx[tris[tri][2]];
        double tmin = 1;
        double q = tet_quality(t0, t1, t2, ex[1]);
        if (q < tmin) tmin = q;
        q = tet_quality(t1, t0, t2, ex[0]);
        if (q < tmin) tmin = q;
        cached |= (1ull << tri);
        cache[tri] = tmin;
      }
      double tq = cache[tri];
      if (tq < mesh_minq) mesh_minq = tq;
      if (mesh_minq < 0) break;
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

// swap_quality (swap_qualities.cpp:13-65)
__global__ __launch_bounds__(BLOCK) void swap_quality_f64(
    u32 ne, u32 const* __restrict__ uo, u32 const* __restrict__ ua, u32 const* __restrict__ ud,
    u32 const* __restrict__ ev, u32 const* __restrict__ vt, double const* __restrict__ x,
    double const* __restrict__ eq, u32* __restrict__ cand, double* __restrict__ out, u32* __restrict__ rsz,
    u32* __restrict__ words) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= ne || !cand[i]) return;
  double old_minq = 1;
  for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
    double q = eq[ua[j]];
    if (q < old_minq) old_minq = q;
  }
  u32 edge_v[2], ring_v[MAX_EDGE_SWAP];
  u32 rs = edge_ring(i, uo, ua, ud, ev, vt, edge_v, ring_v);
  if (rs > MAX_EDGE_SWAP) {
    cand[i] = 0;
    return;
  }
  if (rs < 3) {
    words[W_ERR] = 1u;
    cand[i] = 0;
    return;
  }
  V3 ex[2] = {ld3(x + 3 * (size_t) edge_v[0]), ld3(x + 3 * (size_t) edge_v[1])};
  V3 rx[MAX_EDGE_SWAP];
  for (u32 j = 0; j < rs; ++j) rx[j] = ld3(x + 3 * (size_t) ring_v[j]);
  double sq;
  choose_edge_swap(rs, ex, rx, &sq);
  if (sq > old_minq) {
    out[i] = sq;
    rsz[i] = rs;
    words[W_ANY] = 1u;
  } else {
    cand[i] = 0;
  }
}

// count_swap_ents (edge_swap.cpp:375-386)
static inline __host__ __device__ u32 count_swap_ents(u32 rs, u32 d, u32 nint, u32 tpm) {
  return d == 1 ? nint : (d == 2 ? tpm + 2 * nint : 2 * tpm);
}

// swap_cavity (swap_topology.cpp:46-75) with get_swap_ents
// (edge_swap.cpp:282-373): recompute the ring and choice, emit products
__global__ __launch_bounds__(BLOCK) void swap_cavity_f64(
    u32 ne, u32 d, u32 const* __restrict__ indset, u32 const* __restrict__ goff, u32 const* __restrict__ uo,
    u32 const* __restrict__ ua, u32 const* __restrict__ ud, u32 const* __restrict__ ev,
    u32 const* __restrict__ vt, double const* __restrict__ x, u32* __restrict__ out) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= ne || !indset[i]) return;
  u32 edge_v[2], ring_v[MAX_EDGE_SWAP];
  u32 rs = edge_ring(i, uo, ua, ud, ev, vt, edge_v, ring_v);
  if (rs < 3 || rs > MAX_EDGE_SWAP) return;  // cannot happen for an indset edge
  V3 ex[2] = {ld3(x + 3 * (size_t) edge_v[0]), ld3(x + 3 * (size_t) edge_v[1])};
  V3 rx[MAX_EDGE_SWAP];
  for (u32 j = 0; j < rs; ++j) rx[j] = ld3(x + 3 * (size_t) ring_v[j]);
  double sq;
  u32 code = choose_edge_swap(rs, ex, rx, &sq);
  u32 tpm = k_sw_sizes[rs];
  unsigned char const* mesh = k_sw_meshes + k_sw_mesh_off[rs] + code * tpm;
  unsigned char const(*tris)[3] = k_sw_tris + k_sw_tri_off[rs];
  u32 nint = k_sw_nint[rs];
  unsigned char const* ie = k_sw_int + k_sw_int_off[rs] + code * nint * 2;
  u32* o = out + (size_t) goff[i] * (d + 1);
  if (d == 3) {
    for (u32 t = 0; t < tpm; ++t) {
      u32 tv[4];
      for (u32 j = 0; j < 3; ++j) tv[j] = ring_v[tris[mesh[t]][j]];
      for (u32 j = 0; j < 2; ++j) {
        tv[3] = edge_v[1 - j];
        for (u32 k = 0; k < 4; ++k) *o++ = tv[k];
        u32 tmp = tv[0];
        tv[0] = tv[1];
        tv[1] = tmp;
      }
    }
  } else if (d == 1) {
    for (u32 k = 0; k < nint; ++k) {
      o[2 * k] = ring_v[ie[2 * k]];
      o[2 * k + 1] = ring_v[ie[2 * k + 1]];
    }
  } else {
    for (u32 k = 0; k < nint; ++k)
      for (u32 j = 0; j < 2; ++j) {
        o[k * 6 + j * 3 + 0] = ring_v[ie[2 * k]];
        o[k * 6 + j * 3 + 1] = ring_v[ie[2 * k + 1]];
        o[k * 6 + j * 3 + 2] = edge_v[j];
      }
    o += nint * 6;
    for (u32 t = 0; t < tpm; ++t)
      for (u32 j = 0; j < 3; ++j) o[t * 3 + j] = ring_v[tris[mesh[t]][j]];
  }
}

// swap_common + swap_interior + swap_ents (swap.cpp:27-136)
static u32 swap_common(Mesh& m, Buf<u32>& cand) {
  u32 ne = count(m, 1);
  if (!max_u32(cand.get(), ne)) return 0;
  unmark_boundary(m, 1, cand.get());
  if (!max_u32(cand.get(), ne)) return 0;
  make_swap_tables();
  Up const& u = ask_up(m, 1, 3);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vt = ask_down(m, 3, 0);
  Buf<double> eq = mesh_qualities(m);
  Buf<double> q(ne);
  Buf<u32> rsz = filled(ne, 0);
  zero_words(0, 2);
  {
    ProfScope ps("swap_quality_f64", (double) ne * 24 + (double) u.adj.size() * 8 + (double) count(m, 3) * 24 + (double) count(m, 0) * 24);
    hipLaunchKernelGGL(swap_quality_f64, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, u.offsets.get(),
                       u.adj.get(), u.dirs.get(), ev, vt, coords(m), eq.get(), cand.get(), q.get(), rsz.get(),
                       scratch_words());
    check_launch("swap_quality_f64");
  }
  eq.release();
  u32 words[2];
  to_host(words, scratch_words(), 2);
  OHX_REQUIRE(!words[W_ERR], "find_edge_ring: ring does not close or has fewer than 3 tets");
  if (!words[W_ANY]) return 0;
  Buf<u32> indset = find_indset(m, 1, cand.get(), q.get());
  cand.release();
  q.release();
  u32 total = sum_u32(indset.get(), ne);
  u32 nv = count(m, 0);
  Mesh mo(m.dim, m.full);
  set_ents(mo, 0, nv, Buf<u32>());
  for (auto& tp : m.tags[0]) {  // copy_tags
    Tag& t = *tp;
    if (t.type == TagType::F64)
      add_tag_f64(mo, 0, t.name.c_str(), t.ncomps, copy_f64(t.f.get(), nv * t.ncomps), t.transfer);
    else
      add_tag_u32(mo, 0, t.name.c_str(), t.ncomps, copy(t.u.get(), nv * t.ncomps), t.transfer);
  }
  u32 const* is = indset.get();
  u32 const* rs = rsz.get();
  u32 dlo = m.full ? 1 : m.dim;
  for (u32 d = dlo; d <= m.dim; ++d) {
    Buf<u32> cnt(ne);
    {
      u32* c = cnt.get();
      u32 dd = d;
      parallel_for(ne, OHX_LAMBDA(u32 i) {
        c[i] = is[i] ? count_swap_ents(rs[i], dd, k_sw_nint[rs[i]], k_sw_sizes[rs[i]]) : 0u;
      });
    }
    Buf<u32> goff = exscan(cnt.get(), ne);
    cnt.release();
    u32 ngen = read_at(goff.get(), ne);
    u32 vpe = d + 1;
    Buf<u32> gen((size_t) ngen * vpe);
    {
      ProfScope ps("swap_cavity_f64", (double) ne * 12 + (double) ngen * vpe * 4);
      hipLaunchKernelGGL(swap_cavity_f64, dim3(nblocks(ne)), dim3(BLOCK), 0, stream(), ne, d, is, goff.get(),
                         u.offsets.get(), u.adj.get(), u.dirs.get(), ev, vt, coords(m), gen.get());
      check_launch("swap_cavity_f64");
    }
    Buf<u32> old = mark_up(m, 1, d, is);
    u32 n = count(m, d);
    {
      u32* o = old.get();
      parallel_for(n, OHX_LAMBDA(u32 i) { o[i] = !o[i]; });
    }
    Buf<u32> soff = exscan(old.get(), n);
    old.release();
    u32 nsame = read_at(soff.get(), n);
    Buf<u32> vout((size_t) (nsame + ngen) * vpe);
    expand_into_u32(vout.get(), ask_down(m, d, 0), soff.get(), n, vpe);
    d2d(vout.get() + (size_t) nsame * vpe, gen.get(), (size_t) ngen * vpe);
    set_ents(mo, d, nsame + ngen, std::move(vout));
    if (m.full) {
      Prods p;  // setup_swap (inherit.cpp:158-166)
      p.ndoms[0] = n;
      p.offs[0] = soff.get();
      p.ndoms[1] = ne;
      p.offs[1] = goff.get();
      p.finish();
      inherit_class(m, mo, d, p);
    }
  }
  say("swapped", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 swap_slivers(Mesh& m, double good_qual, u32 nlayers) {
  OHX_REQUIRE(m.dim == 3, "swap_slivers needs a 3D mesh");
  Buf<u32> sl = mark_slivers(m, good_qual, nlayers);
  Buf<u32> c = mark_down(m, m.dim, 1, sl.get());
  sl.release();
  return swap_common(m, c);
}

// ---------------------------------------------------------------- adapt
// mesh_adapt (adapt.cpp:19-94)
namespace {
struct Ctx {
  u32 ops = 0, max_ops = 0;
  AdaptStats st;
};
void adapt_summary(Mesh& m) {  // adapt.cpp:22-34
  u32 nel = count(m, m.dim);
  double mq = mesh_min_quality(m);
  u32 ne = count(m, 1);
  Buf<double> s = measure_edges(m);
  double lo = min_f64(s.get(), ne, 1), hi = max_f64(s.get(), ne, 1);
  if (g_verbose)
    std::printf("%10u elements, min quality %.0f%%, metric range %.2f - %.2f\n", nel, mq * 100.0, lo, hi);
}
bool incr_op(Mesh& m, Ctx& c) {  // adapt.cpp:36-45
  if (c.ops > c.max_ops) return false;
  ++c.ops;
  adapt_summary(m);
  return true;
}
enum Kind { REFINE, COARSEN, SWAP, COARSEN_SLIVERS };
u32 run_pass(Mesh& m, Ctx& c, Kind k, double a, double b, u32 n) {
  double tin = count(m, m.dim);
  double t0 = now_s();
  u32 ok = 0;
  switch (k) {
    case REFINE: ok = refine_by_size(m, a); break;
    case COARSEN: ok = coarsen_by_size(m, a, b); break;
    case SWAP: ok = swap_slivers(m, a, n); break;
    case COARSEN_SLIVERS: ok = coarsen_slivers(m, a, n); break;
  }
  sync();
  double t1 = now_s();
  if (ok) {
    c.st.npass += 1;
    c.st.nelems += tin;
    c.st.tpass += t1 - t0;
  }
  return ok;
}
}  // namespace

int adapt(Mesh& m, double size_floor, double good_qual, u32 nlayers, u32 max_ops, AdaptStats* st) {
  Ctx c;
  c.max_ops = max_ops;
  double t0 = now_s();
  int result = 0;
  adapt_summary(m);
  double qf = mesh_min_quality(m);  // satisfy_size (adapt.cpp:47-56)
  if (good_qual < qf) qf = good_qual;
  bool aborted = false;
  while (!aborted && run_pass(m, c, REFINE, qf, 0, 0))
    if (!incr_op(m, c)) aborted = true;
  while (!aborted && run_pass(m, c, COARSEN, qf, size_floor, 0))
    if (!incr_op(m, c)) aborted = true;
  while (!aborted) {  // satisfy_shape (adapt.cpp:58-80)
    double prev = mesh_min_quality(m);
    if (prev >= good_qual) break;
    if (m.dim == 3 && run_pass(m, c, SWAP, good_qual, 0, nlayers)) {
      if (!incr_op(m, c)) aborted = true;
      continue;
    }
    if (run_pass(m, c, COARSEN_SLIVERS, good_qual, 0, nlayers)) {
      if (!incr_op(m, c)) aborted = true;
      continue;
    }
    if (g_verbose) std::printf("warning: shape quality could not be satisfied\n");
    break;
  }
  result = aborted ? -1 : (c.ops > 0 ? 1 : 0);
  sync();
  c.st.ttotal = now_s() - t0;
  if (st) *st = c.st;
  return result;
}

// mesh_warp_to_limit (warp_to_limit.cpp:14-81)
u32 warp_to_limit(Mesh& m, double qual_floor) {
  u32 dim = m.dim;
  u32 nv = count(m, 0), ne = count(m, dim);
  u32 const* ev = ask_down(m, dim, 0);
  double const* x = coords(m);
  Tag* wt = find_tag(m, 0, "warp");
  OHX_REQUIRE(wt && wt->type == TagType::F64 && wt->ncomps == 3, "mesh has no 3-component warp field");
  double const* w = wt->f.get();
  Buf<double> xo((size_t) 3 * nv);
  double* xp = xo.get();
  double f = 1;
  bool hit = false;
  auto axpy = [&](double a) {  // doubles_axpy (doubles.cpp:133-145): a*x + y, unfused
    parallel_for(3 * nv, OHX_LAMBDA(u32 i) { xp[i] = a * w[i] + x[i]; });
  };
  axpy(f);
  while (true) {
    double mq = min_element_quality(dim, ne, ev, xp, nv);
    if (!(mq < qual_floor)) break;
    if (g_verbose) std::printf("quality %f exceeds limit\n", mq);
    hit = true;
    f /= 2;
    axpy(f);
  }
  Buf<double> wo((size_t) 3 * nv);
  double* wp = wo.get();
  double nf = -f;
  parallel_for(3 * nv, OHX_LAMBDA(u32 i) { wp[i] = nf * w[i] + w[i]; });
  remove_tag(m, 0, "coordinates");
  remove_tag(m, 0, "warp");
  add_tag_f64(m, 0, "coordinates", 3, std::move(xo));
  add_tag_f64(m, 0, "warp", 3, std::move(wo));
  return hit ? 0 : 1;
}

}  // namespace ohx
