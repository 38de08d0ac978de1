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
#include <memory>

#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_prims.hpp"
#include "ohx_scan.hpp"

namespace ohx {

thread_local BatchCtx* g_batch = nullptr;
double g_batch_retire_min = 1e6;
double g_batch_handoff = 0.25;

namespace {

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
    // neighbours share a component: one store per component and wave
    wave_flag_at(ch + co(i), indset[i] != 0);
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

// mesh_adapt's control flow for one mesh, resumed from a given phase,
// operation count and quality floor (a batch hands its last components over
// to this when batch passes would mostly copy finished ones)
static void adapt_continue(Mesh& m, int& phase, u32& ops, double qf, double size_floor, double good_qual,
                           u32 nlayers, u32 max_ops, AdaptStats& stats) {
  auto now = [] {
    sync();
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  auto run = [&](int k) {
    double tin = count(m, m.dim);
    double t0 = now();
    u32 ok = 0;
    switch (k) {
      case 0: ok = refine_by_size(m, qf); break;
      case 1: ok = coarsen_by_size(m, qf, size_floor); break;
      case 2: ok = swap_slivers(m, good_qual, nlayers); break;
      case 3: ok = coarsen_slivers(m, good_qual, nlayers); break;
    }
    double t1 = now();
    if (ok) {
      stats.npass += 1;
      stats.nelems += tin;
      stats.tpass += t1 - t0;
    }
    return ok != 0;
  };
  auto incr = [&] {
    if (ops > max_ops) {
      phase = P_ABORTED;
      return false;
    }
    ++ops;
    return true;
  };
  while (phase != P_DONE && phase != P_ABORTED) {
    switch (phase) {
      case P_REFINE:
        if (run(0))
          incr();
        else
          phase = P_COARSEN;
        break;
      case P_COARSEN:
        if (run(1))
          incr();
        else
          phase = P_CHECK;
        break;
      case P_CHECK:
        phase = mesh_min_quality(m) >= good_qual ? P_DONE : (m.dim == 3 ? P_SWAP : P_CSLIVERS);
        break;
      case P_SWAP:
        if (!run(2))
          phase = P_CSLIVERS;
        else if (incr())
          phase = P_CHECK;
        break;
      case P_CSLIVERS:
        if (!run(3))
          phase = P_DONE;
        else if (incr())
          phase = P_CHECK;
        break;
    }
  }
}

void batch_adapt(Mesh& m, BatchCtx& b, double size_floor, double good_qual, u32 nlayers, u32 max_ops, int* results,
                 AdaptStats* st) {
  u32 const nc = b.ncomp;
  auto now = [] {
    sync();
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  double t_start = now();
  AdaptStats stats;
  // per original component
  std::vector<int> phase(nc, P_REFINE);
  std::vector<u32> ops(nc, 0);
  std::vector<double> qf(nc);
  std::vector<std::unique_ptr<Mesh>> retired(nc);
  // the live batch: components still adapting, live[k] = original index of
  // live component k.  A pass costs the whole live batch (every survivor is
  // copied, every adjacency rebuilt), so once most of its elements belong to
  // finished components those are extracted and the rest re-batched.
  Mesh* cur = &m;
  BatchCtx* cb = &b;
  std::unique_ptr<Mesh> own_mesh;
  std::unique_ptr<BatchCtx> own_ctx;
  std::vector<u32> live(nc);
  for (u32 c = 0; c < nc; ++c) live[c] = c;
  std::vector<double> minq, nel;
  g_batch = cb;
  struct Reset {
    ~Reset() { g_batch = nullptr; }
  } reset;
  // satisfy_size (adapt.cpp:47-56): each component's own floor
  batch_quality(*cur, *cb, minq, nel);
  for (u32 c = 0; c < nc; ++c) qf[c] = good_qual < minq[c] ? good_qual : minq[c];
  auto upload_floors = [&] {
    std::vector<double> q(live.size());
    for (size_t k = 0; k < live.size(); ++k) q[k] = qf[live[k]];
    Buf<double> d = to_device(q.data(), q.size());
    d2d(cb->qf.get(), d.get(), q.size());
  };
  upload_floors();
  bool nel_fresh = true;
  auto finished = [&](u32 c) { return phase[c] == P_DONE || phase[c] == P_ABORTED; };
  auto compact = [&] {
    if (!nel_fresh) {
      batch_counts(*cur, *cb, nel);
      nel_fresh = true;
    }
    double all = 0, open = 0;
    u32 nopen = 0;
    for (size_t k = 0; k < live.size(); ++k) {
      all += nel[k];
      if (!finished(live[k])) {
        open += nel[k];
        ++nopen;
      }
    }
    if (nopen == live.size() || all < g_batch_retire_min || open > 0.5 * all) return;
    std::vector<std::unique_ptr<Mesh>> keep;
    std::vector<u32> next;
    for (size_t k = 0; k < live.size(); ++k) {
      auto part = std::make_unique<Mesh>(cur->dim, cur->full);
      batch_extract(*cur, *cb, (u32) k, *part);
      if (finished(live[k])) {
        retired[live[k]] = std::move(part);
      } else {
        keep.push_back(std::move(part));
        next.push_back(live[k]);
      }
    }
    auto nm = std::make_unique<Mesh>(cur->dim, true);
    auto nb = std::make_unique<BatchCtx>();
    if (!keep.empty()) {
      std::vector<Mesh*> ps;
      for (auto& p : keep) ps.push_back(p.get());
      batch_union(ps, *nm, *nb);
    }
    own_mesh = std::move(nm);
    own_ctx = std::move(nb);
    cur = own_mesh.get();
    cb = own_ctx.get();
    g_batch = cb;
    live = next;
    nel_fresh = false;
    if (!live.empty()) upload_floors();
  };
  std::vector<u32> act, ch;
  // one batch pass of kind k on the live components in phase ph
  auto pass = [&](int ph, int k) {
    u32 nl = (u32) live.size();
    act.assign(nl, 0);
    u32 any = 0;
    for (u32 j = 0; j < nl; ++j) {
      act[j] = phase[live[j]] == ph;
      any |= act[j];
    }
    if (!any) return false;
    if (!nel_fresh) {
      batch_counts(*cur, *cb, nel);
      nel_fresh = true;
    }
    {
      Buf<u32> a = to_device(act.data(), nl);
      d2d(cb->active.get(), a.get(), nl);
    }
    fill(cb->changed.get(), nl, 0);
    double t0 = now();
    switch (k) {
      case 0: refine_by_size(*cur, good_qual); break;
      case 1: coarsen_by_size(*cur, good_qual, size_floor); break;
      case 2: swap_slivers(*cur, good_qual, nlayers); break;
      case 3: coarsen_slivers(*cur, good_qual, nlayers); break;
    }
    ch.assign(nl, 0);
    to_host(ch.data(), cb->changed.get(), nl);
    double t1 = now();
    double tin = 0;
    bool changed_any = false;
    for (u32 j = 0; j < nl; ++j)
      if (act[j] && ch[j]) {
        tin += nel[j];
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
  // the batch pass that serves the most elements next; when even that one
  // would mostly copy components it does not change, the rest of every live
  // component's sequence runs one mesh at a time
  static int const kind_phase[4] = {P_REFINE, P_COARSEN, P_SWAP, P_CSLIVERS};
  while (!live.empty()) {
    compact();
    if (live.empty()) break;
    u32 nl = (u32) live.size();
    bool checking = false;
    for (u32 c : live) checking |= phase[c] == P_CHECK;
    if (checking) {  // satisfy_shape (adapt.cpp:58-80)
      batch_quality(*cur, *cb, minq, nel);
      nel_fresh = true;
      for (u32 j = 0; j < nl; ++j) {
        u32 c = live[j];
        if (phase[c] == P_CHECK) phase[c] = minq[j] >= good_qual ? P_DONE : (cur->dim == 3 ? P_SWAP : P_CSLIVERS);
      }
      continue;
    }
    if (!nel_fresh) {
      batch_counts(*cur, *cb, nel);
      nel_fresh = true;
    }
    double total = 0, best = 0;
    int kind = -1;
    for (u32 j = 0; j < nl; ++j) total += nel[j];
    for (int k = 0; k < 4; ++k) {
      double a = 0;
      for (u32 j = 0; j < nl; ++j)
        if (phase[live[j]] == kind_phase[k]) a += nel[j];
      if (a > best) {
        best = a;
        kind = k;
      }
    }
    if (kind < 0) break;  // every live component has finished
    if (nl > 1 && best < g_batch_handoff * total) {
      for (u32 j = 0; j < nl; ++j) {
        u32 c = live[j];
        auto part = std::make_unique<Mesh>(cur->dim, cur->full);
        batch_extract(*cur, *cb, j, *part);
        g_batch = nullptr;
        adapt_continue(*part, phase[c], ops[c], qf[c], size_floor, good_qual, nlayers, max_ops, stats);
        g_batch = cb;
        retired[c] = std::move(part);
      }
      live.clear();
      break;
    }
    pass(kind_phase[kind], kind);
    for (u32 j = 0; j < nl; ++j) {
      if (!act[j]) continue;
      u32 c = live[j];
      switch (kind) {
        case 0:
          if (!ch[j])
            phase[c] = P_COARSEN;
          else
            incr(c);
          break;
        case 1:
          if (!ch[j])
            phase[c] = P_CHECK;
          else
            incr(c);
          break;
        case 2:
          if (!ch[j])
            phase[c] = P_CSLIVERS;
          else if (incr(c))
            phase[c] = P_CHECK;
          break;
        case 3:
          if (!ch[j])
            phase[c] = P_DONE;  // "shape quality could not be satisfied"
          else if (incr(c))
            phase[c] = P_CHECK;
          break;
      }
    }
  }
  for (u32 c = 0; c < nc; ++c) results[c] = phase[c] == P_ABORTED ? -1 : (ops[c] > 0 ? 1 : 0);
  // components retired along the way go back into one batch, in order
  bool any_retired = false;
  for (u32 c = 0; c < nc; ++c) any_retired = any_retired || retired[c] != nullptr;
  if (any_retired) {
    std::vector<std::unique_ptr<Mesh>> tail(nc);
    for (size_t k = 0; k < live.size(); ++k) {
      tail[live[k]] = std::make_unique<Mesh>(cur->dim, cur->full);
      batch_extract(*cur, *cb, (u32) k, *tail[live[k]]);
    }
    std::vector<Mesh*> all(nc);
    for (u32 c = 0; c < nc; ++c) all[c] = retired[c] ? retired[c].get() : tail[c].get();
    Mesh out(m.dim, true);
    BatchCtx nb;
    batch_union(all, out, nb);
    overwrite(m, std::move(out));
    b = std::move(nb);
  }
  stats.ttotal = now() - t_start;
  if (st) *st = stats;
}

}  // namespace ohx
