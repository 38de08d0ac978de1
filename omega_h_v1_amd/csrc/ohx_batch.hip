// Batched adaptation: many independent meshes of uneven sizes adapted in
// one launch sequence.
//
// The meshes are concatenated into one mesh (entities of each dimension
// component after component, class ids shifted so that each component owns
// a range per class dimension).  Every pass keeps the relative order of a
// component's entities (survivors in order, then products in the order of
// their sources), and nothing in a pass couples two components: adjacency,
// independent-set ties (decided by the relative order), the collapse and
// swap choices and the sliver layers all stay inside a component.  What
// mesh_adapt decides per mesh -- the quality floor, whether a pass changed
// the mesh, the operation count, when to move on -- the driver below decides
// per component, running each batch pass only on the components whose own
// sequence is at that pass.  Each component therefore ends exactly as if it
// had been adapted alone.  There is no counterpart in the reference, which
// adapts one mesh per call (adapt.cpp:82-94).
#include <cfloat>
#include <chrono>
#include <cstring>

#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_prims.hpp"
#include "ohx_scan.hpp"

namespace ohx {

BatchCtx* g_batch = nullptr;

namespace {

struct BatchScope {
  explicit BatchScope(BatchCtx* b) { g_batch = b; }
  ~BatchScope() { g_batch = nullptr; }
};

// order-preserving map of doubles to unsigned keys (NaN excluded by callers)
__device__ inline unsigned long long dkey(double q) {
  unsigned long long b = (unsigned long long) __double_as_longlong(q);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
inline double dkey_value(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  double q;
  std::memcpy(&q, &b, 8);
  return q;
}
inline unsigned long long host_dkey(double q) {
  unsigned long long b;
  std::memcpy(&b, &q, 8);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

struct CompOf {
  u32 const* cd;
  u32 const* ci;
  u32 const* off;
  u32 nc;
  __device__ u32 operator()(u32 i) const { return batch_comp(off, nc, cd[i], ci[i]); }
};

CompOf comp_of(Mesh& m, BatchCtx& b, u32 d) {
  return CompOf{u32_tag(m, d, "class_dim"), u32_tag(m, d, "class_id"), b.off.get(), b.ncomp};
}

}  // namespace

void batch_mask(Mesh& m, u32 d, u32* flags) {
  BatchCtx& b = *g_batch;
  CompOf co = comp_of(m, b, d);
  u32 const* act = b.active.get();
  parallel_for(count(m, d), OHX_LAMBDA(u32 i) {
    if (!act[co(i)]) flags[i] = 0;
  });
}

Buf<double> batch_floors(Mesh& m, u32 d) {
  BatchCtx& b = *g_batch;
  u32 n = count(m, d);
  Buf<double> out(n);
  CompOf co = comp_of(m, b, d);
  double const* qf = b.qf.get();
  double* o = out.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { o[i] = qf[co(i)]; });
  return out;
}

void batch_note_changed(Mesh& m, u32 d, u32 const* indset) {
  BatchCtx& b = *g_batch;
  CompOf co = comp_of(m, b, d);
  u32* ch = b.changed.get();
  parallel_for(count(m, d), OHX_LAMBDA(u32 i) {
    if (indset[i]) ch[co(i)] = 1u;  // every writer stores the same value
  });
}

// ------------------------------------------------------------------ union
void batch_union(std::vector<Mesh*> const& parts, Mesh& out, BatchCtx& b) {
  OHX_REQUIRE(!parts.empty(), "batch: no meshes");
  u32 dim = parts[0]->dim;
  u32 np = (u32) parts.size();
  for (Mesh* p : parts) {
    OHX_REQUIRE(p->dim == dim, "batch: meshes of different dimensions");
    for (u32 d = 0; d <= dim; ++d) {
      OHX_REQUIRE(has_dim(*p, d), "batch: every mesh must have all entity dimensions (a full mesh)");
      OHX_REQUIRE(find_tag(*p, d, "class_dim") && find_tag(*p, d, "class_id"),
                  "batch: every mesh must be classified (class_dim and class_id on every dimension)");
    }
  }
  // class-id ranges: component p owns [off[cd][p], off[cd][p + 1]) per class dim
  Buf<u32> maxp1 = filled(4 * np, 0);
  for (u32 p = 0; p < np; ++p)
    for (u32 d = 0; d <= dim; ++d) {
      u32 const* cd = u32_tag(*parts[p], d, "class_dim");
      u32 const* ci = u32_tag(*parts[p], d, "class_id");
      u32* mx = maxp1.get() + 4 * p;
      parallel_for(count(*parts[p], d), OHX_LAMBDA(u32 i) {
        if (cd[i] < 4) atomicMax(&mx[cd[i]], ci[i] + 1);
      });
    }
  std::vector<u32> hm(4 * np);
  to_host(hm.data(), maxp1.get(), 4 * np);
  std::vector<u32> off(4 * (np + 1));
  for (u32 cd = 0; cd < 4; ++cd) {
    unsigned long long acc = 0;
    for (u32 p = 0; p < np; ++p) {
      off[cd * (np + 1) + p] = (u32) acc;
      acc += hm[4 * p + cd];
    }
    OHX_REQUIRE(acc < 0xffffffffull, "batch: class ids overflow 32 bits");
    off[cd * (np + 1) + np] = 0xffffffffu;
  }
  b.ncomp = np;
  b.off = to_device(off.data(), off.size());
  b.active = filled(np, 1);
  b.qf = filled_f64(np, 0.0);
  b.changed = filled(np, 0);

  out.dim = dim;
  out.full = true;
  std::vector<u32> voff(np + 1, 0);
  for (u32 p = 0; p < np; ++p) voff[p + 1] = voff[p] + count(*parts[p], 0);
  for (u32 d = 0; d <= dim; ++d) {
    size_t n = 0;
    for (Mesh* p : parts) n += count(*p, d);
    OHX_REQUIRE(n < 0xffffffffull, "batch: more than 2^32 entities");
    if (d == 0) {
      set_ents(out, 0, (u32) n, Buf<u32>());
    } else {
      Buf<u32> conn(n * (d + 1));
      size_t base = 0;
      for (u32 p = 0; p < np; ++p) {
        u32 k = count(*parts[p], d) * (d + 1);
        u32 const* src = parts[p]->down[d][0].get();
        u32* dst = conn.get() + base;
        u32 vo = voff[p];
        parallel_for(k, OHX_LAMBDA(u32 j) { dst[j] = src[j] + vo; });
        base += k;
      }
      set_ents(out, d, (u32) n, std::move(conn));
    }
    for (auto& tp : parts[0]->tags[d]) {
      Tag& t0 = *tp;
      for (Mesh* p : parts) {
        Tag* t = find_tag(*p, d, t0.name.c_str());
        OHX_REQUIRE(t && t->type == t0.type && t->ncomps == t0.ncomps,
                    "batch: every mesh must carry the same fields");
      }
      bool is_id = t0.name == "class_id";
      if (t0.type == TagType::F64) {
        Buf<double> data(n * t0.ncomps);
        size_t base = 0;
        for (Mesh* p : parts) {
          size_t k = (size_t) count(*p, d) * t0.ncomps;
          d2d(data.get() + base, find_tag(*p, d, t0.name.c_str())->f.get(), k);
          base += k;
        }
        add_tag_f64(out, d, t0.name.c_str(), t0.ncomps, std::move(data), t0.transfer);
      } else {
        Buf<u32> data(n * t0.ncomps);
        size_t base = 0;
        for (u32 p = 0; p < np; ++p) {
          size_t k = (size_t) count(*parts[p], d) * t0.ncomps;
          u32 const* src = find_tag(*parts[p], d, t0.name.c_str())->u.get();
          if (is_id) {
            u32 const* cd = u32_tag(*parts[p], d, "class_dim");
            u32 const* of = b.off.get();
            u32 row = np + 1;
            u32* dst = data.get() + base;
            parallel_for((u32) k, OHX_LAMBDA(u32 j) { dst[j] = src[j] + (cd[j] < 4 ? of[cd[j] * row + p] : 0u); });
          } else {
            d2d(data.get() + base, src, k);
          }
          base += k;
        }
        add_tag_u32(out, d, t0.name.c_str(), t0.ncomps, std::move(data), t0.transfer);
      }
    }
  }
}

// ---------------------------------------------------------------- extract
void batch_extract(Mesh& u, BatchCtx& b, u32 c, Mesh& out) {
  OHX_REQUIRE(c < b.ncomp, "batch: no such component");
  u32 dim = u.dim;
  out.dim = dim;
  out.full = u.full;
  Buf<u32> vmap;
  for (u32 d = 0; d <= dim; ++d) {
    u32 n = count(u, d);
    CompOf co = comp_of(u, b, d);
    Buf<u32> offs(n + 1);
    u32 total = exscan_fn_total(n, [=] __device__(u32 i) { return co(i) == c ? 1u : 0u; }, offs.get());
    u32 const* of = offs.get();
    if (d == 0) {
      vmap = Buf<u32>(n);
      u32* vm = vmap.get();
      parallel_for(n, OHX_LAMBDA(u32 i) { vm[i] = of[i + 1] != of[i] ? of[i] : INVALID; });
      set_ents(out, 0, total, Buf<u32>());
    } else {
      Buf<u32> conn((size_t) total * (d + 1));
      u32 const* src = u.down[d][0].get();
      u32 const* vm = vmap.get();
      u32* dst = conn.get();
      u32 w = d + 1;
      parallel_for(n, OHX_LAMBDA(u32 i) {
        if (of[i + 1] == of[i]) return;
        for (u32 k = 0; k < w; ++k) dst[(size_t) of[i] * w + k] = vm[src[(size_t) i * w + k]];
      });
      set_ents(out, d, total, std::move(conn));
    }
    u32 const* cdt = u32_tag(u, d, "class_dim");
    u32 const* offt = b.off.get();
    u32 row = b.ncomp + 1;
    for (auto& tp : u.tags[d]) {
      Tag& t = *tp;
      u32 w = t.ncomps;
      if (t.type == TagType::F64) {
        Buf<double> data((size_t) total * w);
        double const* src = t.f.get();
        double* dst = data.get();
        parallel_for(n, OHX_LAMBDA(u32 i) {
          if (of[i + 1] == of[i]) return;
          for (u32 k = 0; k < w; ++k) dst[(size_t) of[i] * w + k] = src[(size_t) i * w + k];
        });
        add_tag_f64(out, d, t.name.c_str(), w, std::move(data), t.transfer);
      } else {
        bool is_id = t.name == "class_id";
        Buf<u32> data((size_t) total * w);
        u32 const* src = t.u.get();
        u32* dst = data.get();
        parallel_for(n, OHX_LAMBDA(u32 i) {
          if (of[i + 1] == of[i]) return;
          u32 sub = is_id && cdt[i] < 4 ? offt[cdt[i] * row + c] : 0u;
          for (u32 k = 0; k < w; ++k) dst[(size_t) of[i] * w + k] = src[(size_t) i * w + k] - sub;
        });
        add_tag_u32(out, d, t.name.c_str(), w, std::move(data), t.transfer);
      }
    }
  }
}

// ------------------------------------------------- per-component quality
void batch_quality(Mesh& m, BatchCtx& b, std::vector<double>& minq, std::vector<double>& nelems) {
  u32 dim = m.dim, n = count(m, dim), nc = b.ncomp;
  CompOf co = comp_of(m, b, dim);
  u32 const* ev = ask_down(m, dim, 0);
  double const* x = coords(m);
  double const* mt = metric_of(m);
  Buf<u32> first = filled(nc, INVALID);
  Buf<u32> cnt = filled(nc, 0);
  Buf<unsigned long long> key(nc);
  {
    unsigned long long k0 = host_dkey(DBL_MAX);
    unsigned long long* kp = key.get();
    parallel_for(nc, OHX_LAMBDA(u32 c) { kp[c] = k0; });
  }
  u32* fp = first.get();
  u32* cp = cnt.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    u32 c = co(i);
    atomicMin(&fp[c], i);
    atomicAdd(&cp[c], 1u);
  });
  unsigned long long* kp = key.get();
  // the serial minimum skips each mesh's first element (doubles.cpp:111-118)
  if (dim == 3)
    parallel_for(n, OHX_LAMBDA(u32 i) {
      u32 c = co(i);
      if (i == fp[c]) return;
      double q = elem_q<3>(i, ev, x, mt);
      if (q == q) atomicMin(&kp[c], dkey(q));
    });
  else
    parallel_for(n, OHX_LAMBDA(u32 i) {
      u32 c = co(i);
      if (i == fp[c]) return;
      double q = elem_q<2>(i, ev, x, mt);
      if (q == q) atomicMin(&kp[c], dkey(q));
    });
  std::vector<unsigned long long> hk(nc);
  std::vector<u32> hc(nc);
  to_host(hk.data(), kp, nc);
  to_host(hc.data(), cp, nc);
  minq.resize(nc);
  nelems.resize(nc);
  for (u32 c = 0; c < nc; ++c) {
    minq[c] = dkey_value(hk[c]);
    nelems[c] = hc[c];
  }
}

// per-component element counts
static void batch_counts(Mesh& m, BatchCtx& b, std::vector<double>& nelems) {
  u32 n = count(m, m.dim), nc = b.ncomp;
  CompOf co = comp_of(m, b, m.dim);
  Buf<u32> cnt = filled(nc, 0);
  u32* cp = cnt.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { atomicAdd(&cp[co(i)], 1u); });
  std::vector<u32> hc(nc);
  to_host(hc.data(), cp, nc);
  nelems.assign(hc.begin(), hc.end());
}

// ----------------------------------------------------------------- driver
namespace {
enum Phase { P_REFINE, P_COARSEN, P_CHECK, P_SWAP, P_CSLIVERS, P_DONE, P_ABORTED };
}

void batch_adapt(Mesh& m, BatchCtx& b, double size_floor, double good_qual, u32 nlayers, u32 max_ops, int* results,
                 AdaptStats* st) {
  BatchScope scope(&b);
  u32 nc = b.ncomp;
  auto now = [] {
    sync();
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  double t_start = now();
  AdaptStats stats;
  std::vector<int> phase(nc, P_REFINE);
  std::vector<u32> ops(nc, 0);
  std::vector<double> minq, nel;
  // satisfy_size (adapt.cpp:47-56): each component's own floor
  batch_quality(m, b, minq, nel);
  std::vector<double> qf(nc);
  for (u32 c = 0; c < nc; ++c) qf[c] = good_qual < minq[c] ? good_qual : minq[c];
  {
    Buf<double> q = to_device(qf.data(), nc);
    d2d(b.qf.get(), q.get(), nc);
  }
  std::vector<u32> act(nc), ch(nc);
  bool nel_fresh = true;
  // one batch pass of kind k on the components in phase ph
  auto pass = [&](int ph, int k) {
    u32 any = 0;
    for (u32 c = 0; c < nc; ++c) {
      act[c] = phase[c] == ph;
      any |= act[c];
    }
    if (!any) return false;
    if (!nel_fresh) {
      batch_counts(m, b, nel);
      nel_fresh = true;
    }
    {
      Buf<u32> a = to_device(act.data(), nc);
      d2d(b.active.get(), a.get(), nc);
    }
    fill(b.changed.get(), nc, 0);
    double t0 = now();
    switch (k) {
      case 0: refine_by_size(m, good_qual); break;
      case 1: coarsen_by_size(m, good_qual, size_floor); break;
      case 2: swap_slivers(m, good_qual, nlayers); break;
      case 3: coarsen_slivers(m, good_qual, nlayers); break;
    }
    to_host(ch.data(), b.changed.get(), nc);
    double t1 = now();
    double tin = 0;
    bool changed_any = false;
    for (u32 c = 0; c < nc; ++c)
      if (act[c] && ch[c]) {
        tin += nel[c];
        changed_any = true;
      }
    if (changed_any) {
      stats.npass += 1;
      stats.nelems += tin;
      stats.tpass += t1 - t0;
      nel_fresh = false;
    }
    return true;
  };
  // incr_op_count (adapt.cpp:36-45), per component
  auto incr = [&](u32 c) {
    if (ops[c] > max_ops) {
      phase[c] = P_ABORTED;
      return false;
    }
    ++ops[c];
    return true;
  };
  while (true) {
    if (pass(P_REFINE, 0)) {
      for (u32 c = 0; c < nc; ++c)
        if (act[c]) {
          if (!ch[c])
            phase[c] = P_COARSEN;
          else
            incr(c);
        }
      continue;
    }
    if (pass(P_COARSEN, 1)) {
      for (u32 c = 0; c < nc; ++c)
        if (act[c]) {
          if (!ch[c])
            phase[c] = P_CHECK;
          else
            incr(c);
        }
      continue;
    }
    bool checking = false;
    for (u32 c = 0; c < nc; ++c) checking |= phase[c] == P_CHECK;
    if (checking) {  // satisfy_shape (adapt.cpp:58-80)
      batch_quality(m, b, minq, nel);
      nel_fresh = true;
      for (u32 c = 0; c < nc; ++c)
        if (phase[c] == P_CHECK) phase[c] = minq[c] >= good_qual ? P_DONE : (m.dim == 3 ? P_SWAP : P_CSLIVERS);
      continue;
    }
    if (pass(P_SWAP, 2)) {
      for (u32 c = 0; c < nc; ++c)
        if (act[c]) {
          if (!ch[c])
            phase[c] = P_CSLIVERS;
          else if (incr(c))
            phase[c] = P_CHECK;
        }
      continue;
    }
    if (pass(P_CSLIVERS, 3)) {
      for (u32 c = 0; c < nc; ++c)
        if (act[c]) {
          if (!ch[c])
            phase[c] = P_DONE;  // "shape quality could not be satisfied"
          else if (incr(c))
            phase[c] = P_CHECK;
        }
      continue;
    }
    break;
  }
  for (u32 c = 0; c < nc; ++c) results[c] = phase[c] == P_ABORTED ? -1 : (ops[c] > 0 ? 1 : 0);
  stats.ttotal = now() - t_start;
  if (st) *st = stats;
}

}  // namespace ohx
