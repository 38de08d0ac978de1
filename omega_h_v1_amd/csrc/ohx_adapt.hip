// Refine, coarsen and swap passes, the independent set, the adapt driver
// and warp_to_limit, on a device-resident mesh.  Every per-entity loop is
// a gfx950 kernel; the host only sequences passes and reads counts.
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>

#include "ohx_adapt.hpp"
#include "ohx_geom.hpp"
#include "ohx_prims.hpp"
#include "ohx_scan.hpp"
#include "ohx_tables.hpp"

namespace ohx {

int g_verbose = 0;
// indset.cpp:33 stops after 100 rounds; raising it is an extension
u32 g_indset_max_rounds = 100;
// the most rounds a set that went to the checked loop needed (a set that
// settles in the first unchecked batch needs at most 8)
std::atomic<u32> g_indset_rounds_seen{0};
// FP32 screening with FP64 re-evaluation near every decision (extension)
bool g_mixed = false;

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// scratch word slots (zeroed by a fill kernel before each use)
enum { W_ANY = 0, W_CACHED = 1, W_FLAGS = 8, W_STATS = 40 };
// FP32-mixed statistics: two slotted u64 counters (decisions settled in
// FP32, FP64 re-evaluations) in a buffer of their own, float bits of the
// max |q32 - q64| at scratch word 40
struct ScreenStats {
  unsigned long long* screened;
  unsigned long long* reeval;
  u32* max_err;
};
constexpr size_t STAT_WORDS = (size_t) STAT_SLOTS * STAT_STRIDE;
static unsigned long long* stat_counters() {
  // one buffer per thread, for the life of the thread, zeroed once
  static thread_local unsigned long long* c = nullptr;
  if (!c) {
    c = static_cast<unsigned long long*>(dev_alloc(2 * STAT_WORDS * sizeof(unsigned long long)));
    OHX_CHECK(hipMemsetAsync(c, 0, 2 * STAT_WORDS * sizeof(unsigned long long), stream()));
  }
  return c;
}
static ScreenStats screen_stats() {
  unsigned long long* c = stat_counters();
  return ScreenStats{c, c + STAT_WORDS, scratch_words() + W_STATS};
}

void mixed_stats(double* max_err, unsigned long long* screened, unsigned long long* reeval) {
  std::vector<unsigned long long> c(2 * STAT_WORDS);
  to_host(c.data(), stat_counters(), c.size());
  unsigned long long a = 0, b = 0;
  for (u32 k = 0; k < STAT_SLOTS; ++k) {
    a += c[(size_t) k * STAT_STRIDE];
    b += c[STAT_WORDS + (size_t) k * STAT_STRIDE];
  }
  u32 h = read_at(scratch_words(), W_STATS);
  float e;
  std::memcpy(&e, &h, 4);
  if (max_err) *max_err = e;
  if (screened) *screened = a;
  if (reeval) *reeval = b;
}

void mixed_stats_reset() {
  OHX_CHECK(hipMemsetAsync(stat_counters(), 0, 2 * STAT_WORDS * sizeof(unsigned long long), stream()));
  fill(scratch_words() + W_STATS, 1, 0);
}

// one scratch word, read through the mapped-memory gather (one small kernel
// and a wait, cheaper than a device-to-host copy)
static u32 read_word(u32 k) {
  u32 const* p = scratch_words() + k;
  u32 v = 0;
  read_many(&p, 1, &v);
  return v;
}
// a new value for the tagged "any" word W_ANY (never 0, the cleared value)
static u32 next_any_tag() {
  static thread_local u32 t = 0;
  t = t + 1 ? t + 1 : 1;
  return t;
}

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
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  bool undecided = false;
  if (i < n) {
    u32 s = old[i];
    if (s == UNKNOWN) {
      // one walk over the neighbours: a neighbour in the set decides at
      // once; otherwise a live neighbour with larger goodness (or equal
      // goodness and a lower index) makes this one wait (goodness is read
      // only until the first such neighbour)
      u32 f = offs[i], e = offs[i + 1];
      double const myg = good[i];
      bool wait = false, out = false;
      for (u32 j = f; j < e; ++j) {
        u32 o = adj[j];
        u32 so = old[o];
        if (so == IN_SET) {
          out = true;
          break;
        }
        if (so == NOT_IN_SET || wait) continue;
        double og = good[o];
        if ((myg == og && o < i) || myg < og) wait = true;
      }
      s = out ? (u32) NOT_IN_SET : (wait ? (u32) UNKNOWN : (u32) IN_SET);
      undecided = !out && wait;
    }
    state[i] = s;
  }
  wave_flag(flag, undecided);
}

// The same round over the edges of a 3D mesh with the neighbours taken
// from the tets around each edge (up(1,3), down(3,1)): an edge's
// neighbours are the other edges of its tets, which as a set is the edge
// star star(1,3) (the two other edges of every triangle through the edge
// and the opposite edge of every tet through it); the rule depends only on
// that set, so the star itself, which would need up(1,2), is not built.
__global__ __launch_bounds__(BLOCK) void indset_round_tets(u32 n, u32 const* __restrict__ uo,
                                                           u32 const* __restrict__ ua, u32 const* __restrict__ ud,
                                                           u32 const* __restrict__ er,
                                                           double const* __restrict__ good,
                                                           u32 const* __restrict__ old, u32* __restrict__ state,
                                                           u32* __restrict__ flag) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  bool undecided = false;
  if (i < n) {
    u32 s = old[i];
    if (s == UNKNOWN) {
      double const myg = good[i];
      bool wait = false, out = false;
      for (u32 j = uo[i]; j < uo[i + 1] && !out; ++j) {
        u32 t = ua[j], d = ud[j];
        for (u32 l = 0; l < 6; ++l) {
          if (l == d) continue;
          u32 o = er[6 * (size_t) t + l];
          u32 so = old[o];
          if (so == IN_SET) {
            out = true;
            break;
          }
          if (so == NOT_IN_SET || wait) continue;
          double og = good[o];
          if ((myg == og && o < i) || myg < og) wait = true;
        }
      }
      s = out ? (u32) NOT_IN_SET : (wait ? (u32) UNKNOWN : (u32) IN_SET);
      undecided = !out && wait;
    }
    state[i] = s;
  }
  wave_flag(flag, undecided);
}

// The same round over the vertices, with the neighbours taken from the
// edges (up(0,1), down(1,0)): in a simplicial mesh two vertices share an
// element exactly when they share an edge, so the other ends of a vertex's
// edges are, as a set, its star star(0,dim), which is then not built.
__global__ __launch_bounds__(BLOCK) void indset_round_verts(u32 n, u32 const* __restrict__ uo,
                                                            u32 const* __restrict__ ua, u32 const* __restrict__ ud,
                                                            u32 const* __restrict__ ev,
                                                            double const* __restrict__ good,
                                                            u32 const* __restrict__ old, u32* __restrict__ state,
                                                            u32* __restrict__ flag) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  bool undecided = false;
  if (i < n) {
    u32 s = old[i];
    if (s == UNKNOWN) {
      double const myg = good[i];
      bool wait = false, out = false;
      for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
        u32 o = ev[2 * (size_t) ua[j] + (1 - ud[j])];
        u32 so = old[o];
        if (so == IN_SET) {
          out = true;
          break;
        }
        if (so == NOT_IN_SET || wait) continue;
        double og = good[o];
        if ((myg == og && o < i) || myg < og) wait = true;
      }
      s = out ? (u32) NOT_IN_SET : (wait ? (u32) UNKNOWN : (u32) IN_SET);
      undecided = !out && wait;
    }
    state[i] = s;
  }
  wave_flag(flag, undecided);
}

// rounds == 0: rounds in batches with a host check after each, up to the
// limit (the reference's loop); rounds > 0: exactly that many rounds and no
// host check, *last_flag then points at the device word that is nonzero if
// entities were still undecided after the last round
// start (with start_done rounds already run): the checked loop continues
// from that state instead of from the filter (rounds are a function of the
// state alone, so the result and the round count are those of a run from
// the start)
static Buf<u32> indset_rounds(Mesh& m, u32 ent_dim, u32 const* filter, double const* good, u32 rounds,
                              u32 const** last_flag, bool via_tets = false, Buf<u32>* start = nullptr,
                              u32 start_done = 0) {
  u32 n = count(m, ent_dim);
  // vertices: neighbours through the edges (indset_round_verts)
  bool const via_edges = ent_dim == 0 && m.dim >= 2;
  Graph empty;
  Graph const& g = (via_tets || via_edges) ? empty : ask_star(m, ent_dim, m.dim);
  Up const* ut = via_tets ? &ask_up(m, 1, 3) : via_edges ? &ask_up(m, 0, 1) : nullptr;
  u32 const* er = via_tets ? ask_down(m, 3, 1) : via_edges ? ask_down(m, 1, 0) : nullptr;
  auto round = [&](u32 const* in, u32* outs, u32* fl, double bytes) {
    if (via_tets)
      launch_timed("indset_round", bytes, indset_round_tets, dim3(nblocks(n)), dim3(BLOCK), n, ut->offsets.get(),
                   ut->adj.get(), ut->dirs.get(), er, good, in, outs, fl);
    else if (via_edges)
      launch_timed("indset_round", bytes, indset_round_verts, dim3(nblocks(n)), dim3(BLOCK), n, ut->offsets.get(),
                   ut->adj.get(), ut->dirs.get(), er, good, in, outs, fl);
    else
      launch_timed("indset_round", bytes, indset_round, dim3(nblocks(n)), dim3(BLOCK), n, g.offsets.get(),
                   g.adj.get(), good, in, outs, fl);
  };
  bool const resume = start && *start;
  Buf<u32> a = resume ? std::move(*start) : Buf<u32>(n), b(n);
  constexpr u32 BATCH = 8;
  u32* flags = scratch_words() + W_FLAGS;
  if (!resume) {
    // the states, and the first batch's round flags cleared on the way
    u32* s = a.get();
    parallel_for(n, OHX_LAMBDA(u32 i) {
      s[i] = filter[i] ? UNKNOWN : NOT_IN_SET;
      if (i < BATCH) flags[i] = 0;
    });
  }
  if (!n) {
    if (rounds) {
      fill(flags, 1, 0);
      *last_flag = flags;
    }
    return a;
  }
  // compulsory bytes per round: every state read and written once (the
  // goodness and the star are read only around undecided entities, a count
  // the host does not know, so they are left out: a lower bound)
  double round_bytes = (double) n * 8;
  // rounds are launched in batches; a round after convergence copies the
  // state unchanged, so running past the converged round is harmless
  // rounds per host check: most sets settle within 8 rounds
  u32 const MAX_ROUNDS = g_indset_max_rounds;
  if (rounds) {
    u32 nr = rounds < BATCH ? rounds : BATCH;
    if (n < nr) fill(flags, nr, 0);  // the set-up kernel cleared flags [0, n)
    for (u32 r = 0; r < nr; ++r) {
      round(a.get(), b.get(), flags + r, round_bytes);
      std::swap(a, b);
    }
    *last_flag = flags + nr - 1;
    return a;
  }
  for (u32 done = resume ? start_done : 0; done < MAX_ROUNDS; done += BATCH) {
    // the last batch stops at the limit, as the reference's loop does
    u32 nr = MAX_ROUNDS - done < BATCH ? MAX_ROUNDS - done : BATCH;
    if (done || n < nr) fill(flags, nr, 0);
    for (u32 r = 0; r < nr; ++r) {
      round(a.get(), b.get(), flags + r, round_bytes);
      std::swap(a, b);
    }
    u32 h[BATCH];
    u32 const* fp[BATCH];
    for (u32 r = 0; r < nr; ++r) fp[r] = flags + r;
    read_many(fp, nr, h);
    if (!h[nr - 1]) {
      // h[r]: entities still undecided after round r of this batch
      u32 k = 0;
      while (h[k]) ++k;
      u32 used = done + k + 1, seen = g_indset_rounds_seen.load();
      while (used > seen && !g_indset_rounds_seen.compare_exchange_weak(seen, used)) {
      }
      return a;
    }
  }
  fail("independent set did not converge within the round limit (indset.cpp:107-116)", __FILE__, __LINE__);
  return a;
}

Buf<u32> find_indset(Mesh& m, u32 ent_dim, u32 const* filter, double const* good) {
  return indset_rounds(m, ent_dim, filter, good, 0, nullptr);
}


// a first batch of rounds with no host check; if the set has not settled
// the caller goes on from its state with the checked loop
// (find_indset_resume: same rounds, same result).  The batch never runs
// more rounds than the limit allows, so a set that needs more than the
// limit always reaches the checked loop's failure.
// Four rounds: measured on the 64^3 x 6 anisotropic box (8 cycles), 87 % of the vertex sets and
// 95 % of the edge sets settle within four rounds (most within two), so a
// longer first batch mostly launches rounds that only copy a settled state.
static u32 first_batch_rounds() { return g_indset_max_rounds < 4 ? g_indset_max_rounds : 4; }
static Buf<u32> find_indset_bounded(Mesh& m, u32 ent_dim, u32 const* filter, double const* good,
                                    u32 const** last_flag) {
  return indset_rounds(m, ent_dim, filter, good, first_batch_rounds(), last_flag);
}
// the checked loop from the first batch's state
static Buf<u32> find_indset_resume(Mesh& m, u32 ent_dim, double const* good, Buf<u32>&& state,
                                   bool via_tets = false) {
  return indset_rounds(m, ent_dim, nullptr, good, 0, nullptr, via_tets, &state, first_batch_rounds());
}

// ------------------------------------------------------------- inherit
// Class tags (class_dim, class_id) of products come from the entity they
// were made from (inherit.cpp:14-232); the fused output kernels of each
// pass write them next to the connectivity.

// --------------------------------------------------------------- refine
// Worst child quality of every candidate edge split
// (refine_qualities.cpp:16-82); rejected candidates are cleared.
template <u32 DIM, bool MIXED>
__global__ __launch_bounds__(BLOCK) void refine_quality_f64(
    u32 ns, u32 const* __restrict__ vs, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
    u32 const* __restrict__ ud, u32 const* __restrict__ ve, double const* __restrict__ x,
    double const* __restrict__ mt, double const* __restrict__ eq, double qual_floor_all,
    double const* __restrict__ qfe, u32* __restrict__ cand, double* __restrict__ out, u32* __restrict__ words, u32 tag,
    u32* __restrict__ err, ScreenStats st) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= ns || !cand[i]) return;
  // one floor for the mesh, or its component's in a batch
  double const qual_floor = qfe ? qfe[i] : qual_floor_all;
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
  // the new vertex's metric is the average its field would get on refinement
  Sym msplit;
  if (mt) {
    Sym ma = ld6(mt + 6 * (size_t) vs[2 * i]), mb = ld6(mt + 6 * (size_t) vs[2 * i + 1]);
    for (int k = 0; k < 6; ++k) {
      double s = 0;
      s += ma.m[k];
      s += mb.m[k];
      s /= 2;
      msplit.m[k] = s;
    }
  }
  u32 const* bo = opp(DIM, 0);
  float minq32 = 1.f;
  if (MIXED && !eq) {
    // FP32 screen: a split certainly below the floor is rejected here; one
    // with a child near zero goes to FP64, which also checks q > 0
    bool sure = true;
    for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
      u32 dir = ud[j], e = ua[j];
      u32 const* evs = canon(DIM, 1, 0, dir);
      for (u32 k = 0; k < 2; ++k) {
        u32 const* bv = canon(DIM, DIM - 1, 0, bo[evs[k]]);
        V3 p[VPE];
#pragma unroll
        for (u32 l = 0; l + 1 < VPE; ++l) p[l] = ld3(x + 3 * (size_t) ve[(size_t) e * VPE + bv[l]]);
        p[VPE - 1] = split;
        float q;
        if (mt) {
          Sym pm[VPE];
#pragma unroll
          for (u32 l = 0; l + 1 < VPE; ++l) pm[l] = ld6(mt + 6 * (size_t) ve[(size_t) e * VPE + bv[l]]);
          pm[VPE - 1] = msplit;
          q = quality32<DIM>(p, pm, false);
        } else {
          q = quality32<DIM>(p, nullptr, false);
        }
        if (!(q > SCREEN_MARGIN)) sure = false;
        minq32 = q < minq32 ? q : minq32;
      }
    }
    bool reject = sure && (minq32 < (float) qual_floor - SCREEN_MARGIN);
    wave_count(st.screened, reject);
    if (reject) {
      cand[i] = 0;
      return;
    }
    wave_count(st.reeval, true);
  }
  double minq = 1, old_minq = 1;
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
      double q;
      if (mt) {
        Sym pm[VPE];
#pragma unroll
        for (u32 l = 0; l + 1 < VPE; ++l) pm[l] = ld6(mt + 6 * (size_t) ve[(size_t) e * VPE + bv[l]]);
        pm[VPE - 1] = msplit;
        q = metric_quality<DIM>(p, pm, false);
      } else {
        q = entity_quality(DIM, p);
      }
      if (!(q > 0)) *err = ERR_REFINE_QUALITY;
      if (q < minq) minq = q;
    }
  }
  if (MIXED && !eq && minq == minq) {
    float d = (float) minq - minq32;
    note_max_err(st.max_err, d < 0 ? -d : d);
  }
  if ((minq < qual_floor) || (eq && (minq <= old_minq)))
    cand[i] = 0;
  else {
    out[i] = minq;
    wave_flag_value(words + W_ANY, true, tag);
  }
}

// candidates of refine_by_size (refine.cpp:11-15), fused with measure_edge
__global__ __launch_bounds__(BLOCK) void refine_candidate_f64(u32 n, u32 const* __restrict__ ev,
                                                              double const* __restrict__ x,
                                                              double const* __restrict__ h,
                                                              double const* __restrict__ mt,
                                                              u32* __restrict__ cand, u32* __restrict__ words,
                                                              u32 tag) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 c = edge_size(ev[2 * i], ev[2 * i + 1], x, h, mt) > 1.0 ? 1u : 0u;
  cand[i] = c;
  wave_flag_value(words + W_ANY, c != 0, tag);
}

static Buf<double> refine_qualities(Mesh& m, u32* cand, double qual_floor, bool require_better,
                                    double const* qfe) {
  u32 dim = m.dim;
  u32 ns = count(m, 1);
  u32 const* vs = ask_down(m, 1, 0);
  u32 const* ve = ask_down(m, dim, 0);
  Up const& u = ask_up(m, 1, dim);
  Buf<double> eq;
  if (require_better) eq = mesh_qualities(m);
  Buf<double> out(ns);
  u32 tag = next_any_tag();
  double prof_bytes = (double) ns * 20 + (double) u.adj.size() * 8 + (double) count(m, dim) * (dim + 1) * 4 + (double) count(m, 0) * 24;
  ScreenStats sst = screen_stats();
  // in the mixed mode the FP32 screen is a variant chosen online (it pays
  // only where it rejects enough splits); both give identical results
  static thread_local Tuner tuners[2] = {Tuner("refine_quality_2d_mixed", {0, 1}), Tuner("refine_quality_3d_mixed", {0, 1})};
  auto go = [&](auto kernel, char const* kn, hipEvent_t ea, hipEvent_t eb) {
    if (ea) {
      hipExtLaunchKernelGGL(kernel, dim3(nblocks(ns)), dim3(BLOCK), 0, stream(), ea, eb, 0, ns, vs, u.offsets.get(),
                            u.adj.get(), u.dirs.get(), ve, coords(m), metric_of(m), eq.get(), qual_floor, qfe,
                            cand, out.get(), scratch_words(), tag, dev_error_word(), sst);
      check_launch(kn);
    } else {
      launch_timed(kn, prof_bytes, kernel, dim3(nblocks(ns)), dim3(BLOCK), ns, vs, u.offsets.get(), u.adj.get(),
                   u.dirs.get(), ve, coords(m), metric_of(m), eq.get(), qual_floor, qfe, cand, out.get(),
                   scratch_words(), tag, dev_error_word(), sst);
    }
  };
  auto run = [&](u32 screen, hipEvent_t ea, hipEvent_t eb) {
    if (dim == 3)
      screen ? go(refine_quality_f64<3, true>, "refine_quality_mixed", ea, eb)
             : go(refine_quality_f64<3, false>, "refine_quality_f64", ea, eb);
    else
      screen ? go(refine_quality_f64<2, true>, "refine_quality_mixed", ea, eb)
             : go(refine_quality_f64<2, false>, "refine_quality_f64", ea, eb);
  };
  if (g_mixed)
    tuned(tuners[dim == 3], ns, run);
  else
    run(0, nullptr, nullptr);
  return out;
}

// Up to 8 tag arrays handled by one kernel (FP64 or u32 components).
struct TagList {
  u32 n = 0;
  u32 ncomps[8];
  u32 f64[8];
  void const* src[8];
  void* dst[8];
};

static void tag_list_add(TagList& L, Tag& t, void* dst) {
  OHX_REQUIRE(L.n < 8, "too many vertex tags for one fused kernel");
  L.ncomps[L.n] = t.ncomps;
  L.f64[L.n] = t.type == TagType::F64;
  L.src[L.n] = t.raw();
  L.dst[L.n] = dst;
  ++L.n;
}

// refine_verts + refine_node (refine_common.cpp:29-54, refine_nodal.cpp:10-46):
// every FP64 vertex tag keeps its old values and gets, for each split edge
// in order, the average ((0 + f(v0)) + f(v1)) / 2 of the edge's vertices.
// One thread per output vertex, all tags in the same launch.
__global__ __launch_bounds__(BLOCK) void refine_vertex_tags(u32 nv, u32 ns, u32 const* __restrict__ gen_off,
                                                            u32 const* __restrict__ vs, TagList L) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= nv + ns) return;
  u32 v0, v1;
  bool split = i >= nv;
  if (split) {
    u32 e = i - nv;
    if (gen_off[e] == gen_off[e + 1]) return;
    v0 = vs[2 * e];
    v1 = vs[2 * e + 1];
  } else {
    v0 = v1 = i;
  }
  u32 o = split ? nv + gen_off[i - nv] : i;
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    double const* f = static_cast<double const*>(L.src[t]);
    double* d = static_cast<double*>(L.dst[t]);
    for (u32 c = 0; c < w; ++c) {
      if (split) {
        double sum = 0;
        sum += f[(size_t) v0 * w + c];
        sum += f[(size_t) v1 * w + c];
        sum /= 2;
        d[(size_t) o * w + c] = sum;
      } else {
        d[(size_t) o * w + c] = f[(size_t) i * w + c];
      }
    }
  }
}

static void refine_verts(Mesh& m, Mesh& mo, u32 const* gen_off, u32 nsplit) {
  u32 nv = count(m, 0), ns = count(m, 1);
  set_ents(mo, 0, nv + nsplit, Buf<u32>());
  TagList L;
  std::vector<Tag*> src;
  for (auto& tp : m.tags[0]) {
    Tag& t = *tp;
    if (t.type != TagType::F64) continue;
    Tag& nt = add_tag_f64(mo, 0, t.name.c_str(), t.ncomps, Buf<double>((size_t) (nv + nsplit) * t.ncomps),
                          t.transfer);
    tag_list_add(L, t, nt.f.get());
  }
  if (L.n)
    hipLaunchKernelGGL(refine_vertex_tags, dim3(nblocks(nv + ns)), dim3(BLOCK), 0, stream(), nv, ns, gen_off,
                       ask_down(m, 1, 0), L);
  check_launch("refine_vertex_tags");
}

static u32 prods_per_dom(u32 dom_dim, u32 src_dim, u32 prod_dim) {  // refine_topology.cpp:110-121
  if (prod_dim == 0) return (dom_dim == src_dim) ? 1 : 0;
  u32 base_dim = prod_dim - 1;
  u32 opp_dim = dom_dim - 1 - base_dim;
  if (src_dim < opp_dim) return 0;
  return h_down_deg[src_dim][opp_dim];
}

// Arguments of the fused refine product kernel (one prod dimension).
struct RefineArgs {
  u32 dim, pd;
  u32 nv;
  u32 n[4];                // entity counts
  u32 const* off[4];       // split offsets of domains of each dim (off[1] = edge split offsets)
  u32 const* verts_of[4];  // down(d, 0)
  u32 const* edges_of[4];  // down(d, 1), d >= 2
  u32 const* soff;         // survivor offsets of dim pd (pd >= 1)
  u32 ppd[4];              // products per splitting domain of dim d
  u32 go[4];               // output base of the products of domains of dim d
  u32 seg[5];              // index space: [survivors | doms of d = 1..3]
  u32* out_verts;          // pd >= 1
  u32 const* cd[4];        // class_dim / class_id per dim (may be null)
  u32 const* ci[4];
  u32* out_cd;
  u32* out_ci;
};

// refine_ents (refine_common.cpp:64-129) for one product dimension, fused:
// survivors keep their vertices and class (inherit.cpp:104-138); every
// splitting domain (split_to_domain, splits_to_domains.cpp:13-56: along its
// first split edge) emits its products (refine_domain_entity,
// refine_topology.cpp:26-107: the base opposite each of the edge's
// opposite-dim entities + the new vertex), classified like the domain.
__global__ __launch_bounds__(BLOCK) void refine_products(RefineArgs a) {
  u32 t = xcd_block() * BLOCK + threadIdx.x;
  if (t >= a.seg[4]) return;
  u32 pd = a.pd;
  if (t < a.seg[0]) {
    u32 i = t, o;
    if (pd == 0) {
      o = i;
    } else {
      u32 const* of = a.off[pd];
      if (of[i] != of[i + 1]) return;
      o = a.soff[i];
      for (u32 k = 0; k <= pd; ++k) a.out_verts[(size_t) o * (pd + 1) + k] = a.verts_of[pd][(size_t) i * (pd + 1) + k];
    }
    if (a.out_cd) a.out_cd[o] = a.cd[pd][i];
    if (a.out_ci) a.out_ci[o] = a.ci[pd][i];
    return;
  }
  u32 d = 1;
  while (t >= a.seg[d]) ++d;
  u32 i = t - a.seg[d - 1];
  u32 const* of = a.off[d];
  u32 k0 = of[i];
  if (k0 == of[i + 1]) return;
  u32 ppd = a.ppd[d];
  u32 o = a.go[d] + k0 * ppd;
  if (pd > 0) {
    // the domain's first split edge: its local index and new vertex
    u32 dir = 0, edge = i;
    if (d >= 2) {
      u32 spd = d == 2 ? 3u : 6u;
      for (u32 j = 0; j < spd; ++j) {
        u32 e = a.edges_of[d][(size_t) i * spd + j];
        if (a.off[1][e] != a.off[1][e + 1]) {
          dir = j;
          edge = e;
          break;
        }
      }
    }
    u32 vert = a.nv + a.off[1][edge];
    u32 opp_dim = d - pd;
    u32 const* os = canon(d, 1, opp_dim, dir);
    u32 const* bo = opp(d, opp_dim);
    u32 const* vd = a.verts_of[d] + (size_t) i * (d + 1);
    for (u32 j = 0; j < ppd; ++j) {
      u32 const* bv = canon(d, pd - 1, 0, bo[os[j]]);
      u32* q = a.out_verts + (size_t) (o + j) * (pd + 1);
      for (u32 k = 0; k < pd; ++k) q[k] = vd[bv[k]];
      q[pd] = vert;
    }
  }
  for (u32 j = 0; j < ppd; ++j) {
    if (a.out_cd) a.out_cd[o + j] = a.cd[d][i];
    if (a.out_ci) a.out_ci[o + j] = a.ci[d][i];
  }
}

// offsets of the splitting domains of each dimension d >= 2 (a domain
// splits iff one of its edges does), enqueued without reading their totals
static void refine_dom_offsets(Mesh& m, u32 const* gen_off, Buf<u32>* off) {
  for (u32 d = 2; d <= m.dim; ++d) {
    u32 n = count(m, d);
    u32 const* ed = ask_down(m, d, 1);
    u32 spd = h_down_deg[d][1];
    off[d] = Buf<u32>(n + 1);
    exscan_fn_heavy(n, OHX_LAMBDA(u32 i) {
      for (u32 j = 0; j < spd; ++j) {
        u32 e = ed[(size_t) i * spd + j];
        if (gen_off[e] != gen_off[e + 1]) return 1u;
      }
      return 0u;
    }, off[d].get());
  }
}

// nsplit[d]: splitting domains of dimension d (from the offsets' totals)
static void refine_ents(Mesh& m, Mesh& mo, u32 const* gen_off, Buf<u32>* off, u32 const* nsplit) {
  OHX_REQUIRE(m.full, "refine needs a full-topology mesh");
  u32 dim = m.dim;
  RefineArgs base{};
  base.dim = dim;
  base.nv = count(m, 0);
  base.off[1] = gen_off;  // an edge splits iff it is in the independent set
  for (u32 d = 0; d <= dim; ++d) {
    base.n[d] = count(m, d);
    base.verts_of[d] = d ? ask_down(m, d, 0) : nullptr;
  }
  for (u32 d = 2; d <= dim; ++d) {
    base.edges_of[d] = ask_down(m, d, 1);
    base.off[d] = off[d].get();
  }
  bool has_class = find_tag(m, 0, "class_dim") != nullptr;
  bool has_id = find_tag(m, 0, "class_id") != nullptr;
  if (has_class)
    for (u32 d = 0; d <= dim; ++d) {
      base.cd[d] = u32_tag(m, d, "class_dim");
      base.ci[d] = has_id ? u32_tag(m, d, "class_id") : nullptr;
    }
  for (u32 pd = 0; pd <= dim; ++pd) {
    RefineArgs a = base;
    a.pd = pd;
    u32 nsurv = base.n[pd] - (pd ? nsplit[pd] : 0);
    Buf<u32> soff;
    if (pd) soff = negate_offsets(base.off[pd], base.n[pd]);
    a.soff = soff.get();
    u32 total = nsurv;
    a.seg[0] = base.n[pd];
    for (u32 d = 1; d <= 3; ++d) {
      bool live = d <= dim && d >= (pd > 1 ? pd : 1);
      a.ppd[d] = live ? prods_per_dom(d, 1, pd) : 0;
      a.go[d] = total;
      total += a.ppd[d] * nsplit[d];
      a.seg[d] = a.seg[d - 1] + ((live && a.ppd[d]) ? base.n[d] : 0);
    }
    a.seg[4] = a.seg[3];
    Buf<u32> verts, cdo, cio;
    if (pd) {
      verts = Buf<u32>((size_t) total * (pd + 1));
      a.out_verts = verts.get();
    }
    if (has_class) {
      cdo = Buf<u32>(total);
      a.out_cd = cdo.get();
      if (has_id) {
        cio = Buf<u32>(total);
        a.out_ci = cio.get();
      }
    }
    if (a.seg[4]) {
      hipLaunchKernelGGL(refine_products, dim3(nblocks(a.seg[4])), dim3(BLOCK), 0, stream(), a);
      check_launch("refine_products");
    }
    if (pd) set_ents(mo, pd, total, std::move(verts));
    if (has_class) {
      add_tag_u32(mo, pd, "class_dim", 1, std::move(cdo));
      if (has_id) add_tag_u32(mo, pd, "class_id", 1, std::move(cio));
    }
    if (pd == dim) refine_transfer(m, mo, soff.get(), base.off[dim], a.ppd[dim], a.go[dim], total);
  }
}

static void say(char const* verb, u32 n, u32 d) {
  static char const* const names[4][2] = {{"vertex", "vertices"}, {"edge", "edges"},
                                          {"triangle", "triangles"}, {"tet", "tets"}};
  if (g_verbose) std::printf("%s %10u %s\n", verb, n, names[d][n != 1]);
}

// refine_common (refine_common.cpp:131-190) for edges
// Every adjacency CSR uses 32-bit offsets, as the reference's.  The largest
// one a pass builds is the edge star (3D: 2 uses(1,2) + uses(1,3) = 6F + 6T;
// 2D: 2 uses(1,2) = 6T), and a refine pass can at most double the element
// count; refuse a mesh that could overflow 32 bits instead of wrapping.
static void check_index_space(Mesh& m) {
  size_t t = count(m, m.dim);
  // faces not derived yet: at most 4 per tet
  size_t f = m.dim == 3 ? (has_dim(m, 2) ? (size_t) m.counts[2] : 4 * t) : 0;
  size_t star = m.dim == 3 ? 6 * f + 6 * t : 6 * t;
  if (2 * star >= ((size_t) 1 << 32))
    fail("mesh too large for 32-bit adjacency offsets", __FILE__, __LINE__);
}

static u32 refine_common(Mesh& m, Buf<u32>& cand, double qual_floor, bool require_better,
                         double const* qfe = nullptr) {
  check_index_space(m);
  u32 ns = count(m, 1), dim = m.dim;
  Buf<double> q = refine_qualities(m, cand.get(), qual_floor, require_better, qfe);
  // one synchronisation for the whole decision: a first batch of indset
  // rounds, the split offsets of every dimension, then one read of the
  // convergence flag and the totals (a candidate-free call ends with an
  // empty set and returns 0)
  u32 const* flag = nullptr;
  Buf<u32> indset = find_indset_bounded(m, 1, cand.get(), q.get(), &flag);
  Buf<u32> gen_off(ns + 1);
  Buf<u32> off[4];
  u32 got[4];
  auto offsets_and_totals = [&] {
    exscan(indset.get(), ns, gen_off.get());
    refine_dom_offsets(m, gen_off.get(), off);
    u32 const* ptrs[4] = {flag, gen_off.get() + ns, dim >= 2 ? off[2].get() + count(m, 2) : nullptr,
                          dim >= 3 ? off[3].get() + count(m, 3) : nullptr};
    read_many(ptrs, 4, got);
  };
  offsets_and_totals();
  if (got[0]) {  // not settled in the first batch: the checked loop from there
    indset = find_indset_resume(m, 1, q.get(), std::move(indset));
    flag = nullptr;
    offsets_and_totals();
  }
  if (g_batch) batch_note_changed(m, 1, indset.get());
  cand.release();
  q.release();
  indset.release();
  u32 total = got[1];
  if (!total) return 0;
  u32 nsplit[4] = {0, total, got[2], got[3]};
  Mesh mo(m.dim, m.full);
  refine_verts(m, mo, gen_off.get(), total);
  refine_ents(m, mo, gen_off.get(), off, nsplit);
  say("split", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 refine_by_size(Mesh& m, double qual_floor) {
  u32 ne = count(m, 1);
  double const* h = size_of(m);
  double const* mt = metric_of(m);
  Buf<u32> cand(ne);
  u32 tag = next_any_tag();
  if (ne) {
    double prof_bytes = (double) ne * 12 + (double) count(m, 0) * 32;
    launch_timed("refine_candidate_f64", prof_bytes, refine_candidate_f64, dim3(nblocks(ne)), dim3(BLOCK), ne,
                       ask_down(m, 1, 0), coords(m), h, mt, cand.get(), scratch_words(), tag);
  }
  if (read_word(W_ANY) != tag) return 0;
  if (g_batch) {
    batch_mask(m, 1, cand.get());
    Buf<double> qfe = batch_floors(m, 1);
    return refine_common(m, cand, qual_floor, false, qfe.get());
  }
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

// check_for_exposed_side (check_collapse_class.cpp:57-140), from the
// entities' side: every face (ED = 2) or tet (ED = 3) visits its edges that
// still have a collapse direction and clears the directions it forbids.
// The reference walks each edge's upward adjacency instead; the result is
// an AND over the adjacent entities either way, so the order does not
// matter, and the upward maps up(1,2), up(1,3) and the reflection
// down(3,1) that the walk needs are not built here (a tet's edges come
// from its faces' edges: the edge (a, b) lies in the face opposite a third
// vertex).
template <u32 ED>
__global__ __launch_bounds__(BLOCK) void exposed_sides_k(u32 n, u32 const* __restrict__ vent,
                                                         u32 const* __restrict__ sent,
                                                         u32 const* __restrict__ e_of_f,
                                                         u32 const* __restrict__ v_of_f,
                                                         u32 const* __restrict__ ev, u32 const* __restrict__ ecd,
                                                         u32 const* __restrict__ scd, u32* __restrict__ codes) {
  u32 E = xcd_block() * BLOCK + threadIdx.x;
  if (E >= n) return;
  constexpr u32 VPE = ED + 1, SPE = ED + 1, NE = ED == 3 ? 6 : 3;
  u32 v[VPE], sd[SPE];
#pragma unroll
  for (u32 k = 0; k < SPE; ++k) sd[k] = sent[(size_t) E * SPE + k];
  u32 const cE = ecd[E];
  // a direction is forbidden only through a side classified on a lower
  // dimension than the entity; an entity whose sides all share its class
  // (every interior one) forbids nothing and stops here
  bool exposed = false;
#pragma unroll
  for (u32 k = 0; k < SPE; ++k) exposed = exposed || scd[sd[k]] != cE;
  if (!exposed) return;
#pragma unroll
  for (u32 k = 0; k < VPE; ++k) v[k] = vent[(size_t) E * VPE + k];
  u32 const* so = opp(ED, 0);
#pragma unroll
  for (u32 le = 0; le < NE; ++le) {
    u32 const* lv = canon(ED, 1, 0, le);
    u32 edge;
    if (ED == 2) {
      edge = e_of_f[(size_t) E * 3 + le];
    } else {
      // the face opposite the first tet vertex off the edge holds it
      u32 c3 = (lv[0] != 0 && lv[1] != 0) ? 0u : ((lv[0] != 1 && lv[1] != 1) ? 1u : 2u);
      u32 f = sd[so[c3]];
      u32 const* fv = v_of_f + 3 * (size_t) f;
      u32 va = v[lv[0]], vb = v[lv[1]];
      u32 m = 0;
#pragma unroll
      for (u32 q = 0; q < 3; ++q) {
        u32 a = fv[q], b = fv[(q + 1) % 3];
        if ((a == va && b == vb) || (a == vb && b == va)) m = q;
      }
      edge = e_of_f[3 * (size_t) f + m];
    }
    u32 c = codes[edge];
    if (c == DONT_COLLAPSE) continue;
    u32 mask = 0;
    for (u32 k = 0; k < 2; ++k) {
      if (!collapses(c, k)) continue;
      u32 va = ev[2 * (size_t) edge + (1 - k)];
      u32 l = v[lv[0]] == va ? 0u : 1u;
      if (scd[sd[so[lv[l]]]] != cE) mask |= 1u << k;
    }
    if (mask) atomicAnd(&codes[edge], ~mask);
  }
}

static void check_exposed_sides(Mesh& m, u32 ed, u32* codes) {
  u32 n = count(m, ed);
  if (!n) return;
  u32 const* vent = ask_down(m, ed, 0);
  u32 const* sent = ask_down(m, ed, ed - 1);
  u32 const* e_of_f = ask_down(m, 2, 1);
  u32 const* v_of_f = ask_down(m, 2, 0);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* ecd = u32_tag(m, ed, "class_dim");
  u32 const* scd = u32_tag(m, ed - 1, "class_dim");
  if (ed == 2)
    hipLaunchKernelGGL(exposed_sides_k<2>, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, vent, sent, e_of_f, v_of_f,
                       ev, ecd, scd, codes);
  else
    hipLaunchKernelGGL(exposed_sides_k<3>, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, vent, sent, e_of_f, v_of_f,
                       ev, ecd, scd, codes);
  check_launch("exposed_sides");
}

// coarsen_quality (coarsen_qualities.cpp:12-79)
// G lanes per edge: the elements around the collapsing vertex are split
// over the group and the minima combined with shuffles (a minimum does not
// depend on the order, so the result is the serial loop's).
template <u32 G>
__device__ inline double group_min(double v) {
#pragma unroll
  for (u32 d = G / 2; d >= 1; d >>= 1) {
    double o = __shfl_xor(v, d, G);
    if (o < v) v = o;
  }
  return v;
}

// Modes of the collapse-quality kernel.  CQ_F64: the reference's
// arithmetic.  CQ_SCREEN: an FP32 pass rejects the directions certainly
// below the floor, every other direction is measured again in FP64.
// CQ_CERTIFIED: the FP32 pass settles every direction whose FP32 minimum is
// farther than the margin from the floor -- rejected, or accepted with the
// FP32 minimum as its goodness, marked inexact (bit 2 + direction of the
// code) -- and measures only the rest in FP64.  An inexact goodness is used
// only where every comparison it takes part in is certified (see
// coarsen_certify); CQ_FORCE then measures in FP64 the inexact directions
// of the vertices that need exact values.
enum CqMode : u32 { CQ_F64 = 0, CQ_SCREEN = 1, CQ_CERTIFIED = 2, CQ_FORCE = 3 };
__host__ __device__ inline u32 inexact_bit(u32 d) { return 4u << d; }

template <u32 DIM, bool F32>
__device__ inline double collapse_elem_q(u32 e, u32 dir, u32 gen, V3 const& pg, Sym const& mg,
                                         u32 const* __restrict__ vel, double const* __restrict__ x,
                                         double const* __restrict__ mt, bool* adj) {
  constexpr u32 VPE = DIM + 1;
  u32 const* v = vel + (size_t) e * VPE;
  bool a = false;
#pragma unroll
  for (u32 l = 0; l < VPE; ++l)
    if (v[l] == gen) a = true;
  *adj = a;
  if (a) return 1.0;
  u32 const* bv = canon(DIM, DIM - 1, 0, opp(DIM, 0)[dir]);
  V3 p[VPE];
  p[VPE - 1] = pg;
#pragma unroll
  for (u32 l = 0; l + 1 < VPE; ++l) p[l] = ld3(x + 3 * (size_t) v[bv[l]]);
  if (mt) {
    Sym pm[VPE];
    pm[VPE - 1] = mg;
#pragma unroll
    for (u32 l = 0; l + 1 < VPE; ++l) pm[l] = ld6(mt + 6 * (size_t) v[bv[l]]);
    return F32 ? (double) quality32<DIM>(p, pm, true) : metric_quality<DIM>(p, pm, true);
  }
  return F32 ? (double) quality32<DIM>(p, nullptr, true) : element_quality(DIM, p);
}

struct CqArgs {
  u32 ncand;
  u32 const* list;  // candidate directions 2 e + j, grouped by collapsing vertex
  u32 const* ev;
  u32 const* uo;    // up(0, dim)
  u32 const* ua;
  u32 const* ud;
  u32 const* vel;   // down(dim, 0)
  double const* x;
  double const* mt;
  double const* eq;  // element qualities (require_better) or null
  double qual_floor;
  double const* qfe;  // per-edge floors (batch) or null
  u32* codes;
  double* out;
  u32 const* vmark;  // CQ_FORCE: vertices that need exact goodness
  ScreenStats st;
};

// One collapse direction per group of G lanes.  The directions come
// grouped by collapsing vertex (the list is built from each vertex's edges
// in turn), so the groups of one block share the vertex's element star and
// its vertices in the caches: measured on C3, the edge-ordered form fetched
// about seven times the arrays' size from beyond L2 per launch.  Both
// directions of an edge may be in flight at once, so the code bits are
// cleared (or the inexact bit set) with atomics on the edge's word.
template <u32 DIM, u32 G, u32 MODE>
__global__ __launch_bounds__(BLOCK) void coarsen_quality_f64(CqArgs A) {
  u32 const ncand = A.ncand;
  u32 const* __restrict__ ev = A.ev;
  u32 const* __restrict__ uo = A.uo;
  u32 const* __restrict__ ua = A.ua;
  u32 const* __restrict__ ud = A.ud;
  u32 const* __restrict__ vel = A.vel;
  double const* __restrict__ x = A.x;
  double const* __restrict__ mt = A.mt;
  double const* __restrict__ eq = A.eq;
  u32* __restrict__ codes = A.codes;
  double* __restrict__ out = A.out;
  ScreenStats const st = A.st;
  size_t gid = (size_t) xcd_block() * BLOCK + threadIdx.x;
  size_t slot = gid / G;
  u32 lane = (u32) (gid % G);
  if (slot >= ncand) return;  // whole groups leave together
  u32 const ed = A.list[slot];
  u32 const i = ed >> 1, j = ed & 1u;
  u32 const col = ev[2 * (size_t) i + j], gen = ev[2 * (size_t) i + (1 - j)];
  if (MODE == CQ_FORCE && !((codes[i] & inexact_bit(j)) && A.vmark[col])) return;
  // one floor for the mesh, or its component's in a batch
  double const qual_floor = A.qfe ? A.qfe[i] : A.qual_floor;
  bool const f32 = (MODE == CQ_SCREEN || MODE == CQ_CERTIFIED) && !eq;
  V3 const pg = ld3(x + 3 * (size_t) gen);
  Sym mg;
  if (mt) mg = ld6(mt + 6 * (size_t) gen);
  double minq32 = 1;
  if (f32) {
    // FP32: a direction certainly below the floor is rejected at its first
    // element below it; CQ_CERTIFIED also accepts one certainly above
    bool nan = false;
    double const lo = qual_floor - (double) SCREEN_MARGIN;
    for (u32 k = uo[col] + lane; k < uo[col + 1]; k += G) {
      bool adj;
      double q = collapse_elem_q<DIM, true>(ua[k], ud[k], gen, pg, mg, vel, x, mt, &adj);
      if (adj) continue;
      if (!(q == q)) nan = true;
      if (q < minq32) minq32 = q;
      if (minq32 < lo) break;
    }
    minq32 = group_min<G>(minq32);
#pragma unroll
    for (u32 d = G / 2; d >= 1; d >>= 1) nan = nan || __shfl_xor((int) nan, d, G);
    bool reject = !nan && minq32 < lo;
    bool accept = MODE == CQ_CERTIFIED && !nan && minq32 > qual_floor + (double) SCREEN_MARGIN;
    if (lane == 0) wave_count(st.screened, reject || accept);
    if (reject) {
      if (lane == 0) atomicAnd(&codes[i], ~((1u << j) | inexact_bit(j)));
      return;
    }
    if (accept) {
      if (lane == 0) {
        out[2 * (size_t) i + j] = minq32;
        atomicOr(&codes[i], inexact_bit(j));
      }
      return;
    }
    if (lane == 0) wave_count(st.reeval, true);
  }
  if (MODE == CQ_FORCE && lane == 0) wave_count(st.reeval, true);
  double minq = 1, old_minq = 1;
  for (u32 k = uo[col] + lane; k < uo[col + 1]; k += G) {
    u32 e = ua[k];
    if (eq && eq[e] < old_minq) old_minq = eq[e];
    bool adj;
    double q = collapse_elem_q<DIM, false>(e, ud[k], gen, pg, mg, vel, x, mt, &adj);
    if (adj) continue;
    if (q < minq) minq = q;
    // below the floor the direction is rejected whatever the remaining
    // elements give (the minimum only falls), so the lane stops; the
    // screened kernels keep the full minimum for their error statistic
    if (MODE == CQ_F64 && minq < qual_floor) break;
  }
  minq = group_min<G>(minq);
  if (eq) old_minq = group_min<G>(old_minq);
  if (lane != 0) return;
  if ((f32 || MODE == CQ_FORCE) && minq == minq) {
    double d = MODE == CQ_FORCE ? minq - out[2 * (size_t) i + j] : minq - minq32;
    note_max_err(st.max_err, (float) (d < 0 ? -d : d));
  }
  if ((minq < qual_floor) || (eq && (minq <= old_minq))) {
    atomicAnd(&codes[i], ~((1u << j) | inexact_bit(j)));
  } else {
    out[2 * (size_t) i + j] = minq;
    if (MODE == CQ_FORCE) atomicAnd(&codes[i], ~inexact_bit(j));
  }
}

// collapse_to_vert (collapses_to_verts.cpp:13-42): whether each vertex
// collapses and the best goodness of its collapse directions; inex (when
// given) says whether that goodness is an FP32 value
__global__ __launch_bounds__(BLOCK) void collapse_to_vert_k(u32 nv, u32 const* __restrict__ uo,
                                                            u32 const* __restrict__ ua, u32 const* __restrict__ ud,
                                                            u32 const* __restrict__ cd, double const* __restrict__ q,
                                                            u32 const* __restrict__ vo, u32* __restrict__ cand,
                                                            double* __restrict__ vq, u32* __restrict__ inex) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= nv) return;
  double maxq = 0;
  u32 is = 0, mi = 0;
  // vo: the candidate directions' offsets per vertex; a vertex that had
  // none has none left and reads nothing
  u32 const jb = (vo && vo[i] == vo[i + 1]) ? uo[i + 1] : uo[i];
  for (u32 j = jb; j < uo[i + 1]; ++j) {
    u32 e = ua[j], d = ud[j];
    u32 c = cd[e];
    if (!collapses(c, d)) continue;
    double qq = q[2 * (size_t) e + d];
    if (!is) {
      is = 1;
      maxq = qq;
      mi = (c & inexact_bit(d)) ? 1u : 0u;
    } else if (qq > maxq) {
      maxq = qq;
      mi = (c & inexact_bit(d)) ? 1u : 0u;
    }
  }
  cand[i] = is;
  vq[i] = is ? maxq : 0.0;
  if (inex) inex[i] = mi;
}

// The comparisons an FP32 goodness takes part in, certified or not: a
// collapsing vertex needs exact goodness (mark[v] = 1) if one of its
// directions within the margin of its best is inexact while another is
// also within it (the destination is the first direction equal to the
// best, collapses_to_verts.cpp:86-101), or if a neighbouring candidate's
// best lies within the margin of its own and either is inexact (the
// independent set compares them with == and <, indset.cpp:72-83).  FP32
// values are within half the margin of their FP64 values, so every
// comparison left unmarked has the FP64 outcome.
__global__ __launch_bounds__(BLOCK) void coarsen_certify(u32 nv, u32 const* __restrict__ uo,
                                                         u32 const* __restrict__ ua, u32 const* __restrict__ ud,
                                                         u32 const* __restrict__ cd, double const* __restrict__ q,
                                                         u32 const* __restrict__ ev,
                                                         u32 const* __restrict__ cand, double const* __restrict__ vq,
                                                         u32 const* __restrict__ inex, u32* __restrict__ mark) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= nv) return;
  u32 mk = 0;
  if (cand[i]) {
    double const m = (double) SCREEN_MARGIN;
    double const best = vq[i];
    u32 near = 0, near_inex = 0;
    for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
      u32 e = ua[j], d = ud[j];
      u32 c = cd[e];
      if (!collapses(c, d)) continue;
      double qq = q[2 * (size_t) e + d];
      if (!(qq < best - m)) {
        ++near;
        if (c & inexact_bit(d)) near_inex = 1;
      }
    }
    if (near > 1 && near_inex) mk = 1;
    // the neighbours the independent set compares: the other ends of the
    // vertex's edges (the vertex star, as a set)
    u32 const mine = inex[i];
    for (u32 k = uo[i]; !mk && k < uo[i + 1]; ++k) {
      u32 o = ev[2 * (size_t) ua[k] + (1 - ud[k])];
      if (!cand[o]) continue;
      double dd = vq[o] - best;
      if (!(dd > m || dd < -m) && (mine || inex[o])) mk = 1;
    }
  }
  mark[i] = mk;
}

// coarsen_size_code (coarsen.cpp:13-21), fused with measure_edge
__global__ __launch_bounds__(BLOCK) void coarsen_size_code_f64(u32 n, u32 const* __restrict__ ev,
                                                               double const* __restrict__ x,
                                                               double const* __restrict__ h,
                                                               double const* __restrict__ mt, double floor_,
                                                               u32* __restrict__ codes, u32* __restrict__ words,
                                                               u32 tag) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  double sz = edge_size(ev[2 * i], ev[2 * i + 1], x, h, mt);
  u32 c = (sz < floor_) ? (u32) COLLAPSE_BOTH : (u32) DONT_COLLAPSE;
  codes[i] = c;
  wave_flag_value(words + W_ANY, c != 0, tag);
}

// size codes carried to a collapse pass's new mesh: threads [0, ne_old)
// give a surviving edge its old code; the next ngen measure a product edge
__global__ __launch_bounds__(BLOCK) void carry_size_codes_k(u32 ne_old, u32 ngen, u32 nsame,
                                                            u32 const* __restrict__ soff,
                                                            u32 const* __restrict__ old_codes,
                                                            u32 const* __restrict__ ev, double const* __restrict__ x,
                                                            double const* __restrict__ h,
                                                            double const* __restrict__ mt, double floor_,
                                                            u32* __restrict__ codes) {
  u32 t = xcd_block() * BLOCK + threadIdx.x;
  if (t < ne_old) {
    u32 o = soff[t];
    if (o != soff[t + 1]) codes[o] = old_codes[t];
  } else if (t - ne_old < ngen) {
    u32 i = nsame + (t - ne_old);
    double sz = edge_size(ev[2 * i], ev[2 * i + 1], x, h, mt);
    codes[i] = (sz < floor_) ? (u32) COLLAPSE_BOTH : (u32) DONT_COLLAPSE;
  }
}

// tags_subset (subset.cpp:15-43): every tag of the surviving vertices,
// one thread per old vertex, all tags in one launch
__global__ __launch_bounds__(BLOCK) void subset_tags(u32 n, u32 const* __restrict__ soff, TagList L) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  u32 o = soff[i];
  if (o == soff[i + 1]) return;
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    if (L.f64[t]) {
      double const* f = static_cast<double const*>(L.src[t]);
      double* dd = static_cast<double*>(L.dst[t]);
      for (u32 c = 0; c < w; ++c) dd[(size_t) o * w + c] = f[(size_t) i * w + c];
    } else {
      u32 const* f = static_cast<u32 const*>(L.src[t]);
      u32* dd = static_cast<u32*>(L.dst[t]);
      for (u32 c = 0; c < w; ++c) dd[(size_t) o * w + c] = f[(size_t) i * w + c];
    }
  }
}

// The new entities' sides of dimension l (0 < l < d), carried from the old
// map instead of reflected (reflect_down.cpp:13-81 gives the same map: an
// entity's lows are unique in a conforming mesh).  A survivor keeps its
// sides, renumbered.  A product is the old entity with its collapsing vertex
// replaced by the target; a side without that vertex survives, and one with
// it is either rebuilt on the target itself (a product of dimension l) or
// fused into the side that already joins the target to the same vertices.
struct CoarsenSides {
  u32* out;            // new entities x sides (null: not built)
  u32 const* old;      // the old entities' sides
  u32 const* soff;     // old side -> its index among the surviving sides
  u32 const* egoff;    // old side -> its index among the product sides
  u32 const* fused_to; // fused old side -> the old side it merges into
  u32 nsame, l;        // surviving sides, their dimension
  u32 nold;            // old sides
  u32* err;
};

struct CoarsenArgs {
  u32 d, n, nsame;
  u32 const* vent;
  u32 const* esoff;  // survivor offsets
  u32 const* egoff;  // generated offsets
  u32 const* gv;     // target vertex of generating entities
  u32 const* gd;     // local index of their collapsing vertex
  u32 const* vsoff;  // vertex renumbering (survivor offsets of vertices)
  u32* out;
  u32 const* cd;
  u32 const* ci;
  u32* out_cd;
  u32* out_ci;
  CoarsenSides sides, edges;  // down(d, d-1); the tets' edges (sliver passes)
};

// sides of dimension l of a d-simplex
__host__ __device__ inline u32 simplex_sides(u32 d, u32 l) { return d == 3 ? (l == 1 ? 6u : 4u) : 3u; }

// the new id of side c (dimension s.l) of a product: the old entity's
// sides in row, its collapsing vertex at old local index gd; the product's
// local vertices are the base opposite gd in canonical order, then the
// target in gd's place
__device__ inline u32 coarsen_product_side(u32 d, u32 c, u32 gd, u32 const* row, CoarsenSides const& s) {
  u32 const* bv = canon(d, d - 1, 0, opp(d, 0)[gd]);
  u32 const* cs = canon(d, s.l, 0, c);
  u32 mask = 0;
  bool on_target = false;
  for (u32 q = 0; q <= s.l; ++q) {
    u32 const nl = cs[q];
    on_target = on_target || nl == d;
    mask |= 1u << (nl < d ? bv[nl] : gd);
  }
  u32 k = 0;
  for (u32 j = 0; j < simplex_sides(d, s.l); ++j) {
    u32 const* os = canon(d, s.l, 0, j);
    u32 mj = 0;
    for (u32 q = 0; q <= s.l; ++q) mj |= 1u << os[q];
    if (mj == mask) k = j;
  }
  u32 const side = row[k];
  if (!on_target) return s.soff[side];
  if (s.egoff[side] != s.egoff[side + 1]) return s.nsame + s.egoff[side];
  u32 const into = s.fused_to[side];
  if (into < s.nold) return s.soff[into];
  *s.err = ERR_REFLECT_NOT_FOUND;  // a side on the target that is neither rebuilt nor fused
  return INVALID;
}

// concat_verts_of_ents (inherit.cpp:234-255) + coarsen_elem
// (coarsen_topology.cpp:9-44) + remap_conn (coarsen_common.cpp:29-33) +
// setup_coarsen class inheritance (inherit.cpp:140-156), fused: survivors
// first, then the entities rebuilt on the collapse target (the base
// opposite the collapsing vertex, then the target), vertex ids renumbered.
__global__ __launch_bounds__(BLOCK) void coarsen_outputs(CoarsenArgs a) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= a.n) return;
  u32 d = a.d, vpe = d + 1;
  u32 const* v = a.vent + (size_t) i * vpe;
  u32 o;
  if (a.esoff[i] != a.esoff[i + 1]) {
    o = a.esoff[i];
    for (u32 j = 0; j < vpe; ++j) a.out[(size_t) o * vpe + j] = a.vsoff[v[j]];
    for (CoarsenSides const* sp : {&a.sides, &a.edges}) {
      CoarsenSides const& s = *sp;
      if (!s.out) continue;
      u32 const ns = simplex_sides(d, s.l);
      for (u32 c = 0; c < ns; ++c) s.out[(size_t) o * ns + c] = s.soff[s.old[(size_t) i * ns + c]];
    }
  } else if (a.egoff[i] != a.egoff[i + 1]) {
    o = a.nsame + a.egoff[i];
    u32 const* bv = canon(d, d - 1, 0, opp(d, 0)[a.gd[i]]);
    for (u32 j = 0; j + 1 < vpe; ++j) a.out[(size_t) o * vpe + j] = a.vsoff[v[bv[j]]];
    a.out[(size_t) o * vpe + vpe - 1] = a.vsoff[a.gv[i]];
    for (CoarsenSides const* sp : {&a.sides, &a.edges}) {
      CoarsenSides const& s = *sp;
      if (!s.out) continue;
      u32 const ns = simplex_sides(d, s.l);
      for (u32 c = 0; c < ns; ++c)
        s.out[(size_t) o * ns + c] = coarsen_product_side(d, c, a.gd[i], s.old + (size_t) i * ns, s);
    }
  } else {
    return;
  }
  if (a.out_cd) a.out_cd[o] = a.cd[i];
  if (a.out_ci) a.out_ci[o] = a.ci[i];
}

static char const* cq_name(u32 mode) {
  return mode == CQ_F64 ? "coarsen_quality_f64"
                        : mode == CQ_SCREEN ? "coarsen_quality_mixed"
                                            : mode == CQ_CERTIFIED ? "coarsen_quality_certified" : "coarsen_quality_force";
}

template <u32 DIM, u32 MODE>
static void cq_launch(u32 g, CqArgs const& a, double bytes, hipEvent_t ea, hipEvent_t eb) {
  dim3 grid(nblocks((size_t) a.ncand * g));
  char const* kn = cq_name(MODE);
  auto one = [&](auto kernel) {
    if (ea) {
      hipExtLaunchKernelGGL(kernel, grid, dim3(BLOCK), 0, stream(), ea, eb, 0, a);
      check_launch(kn);
    } else {
      launch_timed(kn, bytes, kernel, grid, dim3(BLOCK), a);
    }
  };
  if (g == 1) one(coarsen_quality_f64<DIM, 1, MODE>);
  else if (g == 2) one(coarsen_quality_f64<DIM, 2, MODE>);
  else if (g == 4) one(coarsen_quality_f64<DIM, 4, MODE>);
  else one(coarsen_quality_f64<DIM, 8, MODE>);
}

template <u32 DIM>
static void cq_launch_mode(u32 mode, u32 g, CqArgs const& a, double bytes, hipEvent_t ea, hipEvent_t eb) {
  switch (mode) {
    case CQ_F64: cq_launch<DIM, CQ_F64>(g, a, bytes, ea, eb); break;
    case CQ_SCREEN: cq_launch<DIM, CQ_SCREEN>(g, a, bytes, ea, eb); break;
    case CQ_CERTIFIED: cq_launch<DIM, CQ_CERTIFIED>(g, a, bytes, ea, eb); break;
    default: cq_launch<DIM, CQ_FORCE>(g, a, bytes, ea, eb); break;
  }
}

// ------------------------------------------------ maps carried across a pass
// A pass numbers the surviving entities of every dimension first, in their
// old relative order and vertex order, and appends its products; the lows
// of a surviving entity survive too.  So the new mesh's upward adjacency of
// a (low, high) pair follows from the old one (up_from_down.cpp:9-47 gives
// the same lists: per low, its uses in increasing use index high * deg +
// local index): a surviving low's list is its old list without the removed
// highs, renumbered, followed by its uses in product highs; a product low
// has product uses only.  Only the product uses (a few per cavity) take
// atomics and a sort.
struct UpPatch {
  u32 deg, nl_old, nsame_l, ngen_l, nsame_h, ngen_h;
  u32 nuse;           // (nsame_h + ngen_h) * deg: writes stay below it
  u32 const* uo;      // old up(l,h): offsets, highs, local indices
  u32 const* ua;
  u32 const* ud;
  u32 const* soff_l;  // old low / high -> index among the survivors
  u32 const* soff_h;
  u32 const* dn;      // the new down(h,l)
  u32* cnt;           // per new low: its uses (surviving, then + product)
  u32* cnt_s;         // per new low: its uses in surviving highs
  u32* full;          // per new low: 1 if it kept every old use (null: not kept)
  u32* slot;          // per product use: its place in its low's list
  u32 const* off;     // the new offsets
  u32* key;           // product use indices at their unsorted places
  u32* adj;
  u32* dirs;
};

// G lanes per old low (a group): a surviving low counts its surviving
// highs, G uses at a time; groups past nl_old start the product lows at
// zero (lane 0 of each)
template <u32 G>
__global__ __launch_bounds__(BLOCK) void up_patch_count_k(UpPatch a) {
  size_t gid = (size_t) xcd_block() * BLOCK + threadIdx.x;
  u32 const t = (u32) (gid / G), lane = (u32) (gid % G);
  if (t < a.nl_old) {
    u32 n = a.soff_l[t];
    if (n == a.soff_l[t + 1]) return;
    u32 c = 0;
    for (u32 k = a.uo[t] + lane; k < a.uo[t + 1]; k += G) {
      u32 h = a.ua[k];
      c += a.soff_h[h] != a.soff_h[h + 1] ? 1u : 0u;
    }
#pragma unroll
    for (u32 d = G / 2; d >= 1; d >>= 1) c += (u32) __shfl_xor((int) c, (int) d, (int) G);
    if (lane == 0) {
      a.cnt[n] = c;
      a.cnt_s[n] = c;
      if (a.full) a.full[n] = c == a.uo[t + 1] - a.uo[t] ? 1u : 0u;
    }
  } else if (t - a.nl_old < a.ngen_l && lane == 0) {
    u32 n = a.nsame_l + (t - a.nl_old);
    a.cnt[n] = 0;
    a.cnt_s[n] = 0;
    if (a.full) a.full[n] = 0;
  }
}

// a product use's low, or INVALID where the product's lows were not found
// (reported where they were built)
__device__ inline u32 up_patch_low(UpPatch const& a, u32 u) {
  u32 e = a.dn[(size_t) a.nsame_h * a.deg + u];
  return e < a.nsame_l + a.ngen_l ? e : INVALID;
}

// each product use takes a place after its low's surviving uses
__global__ __launch_bounds__(BLOCK) void up_patch_slot_k(UpPatch a) {
  u32 u = xcd_block() * BLOCK + threadIdx.x;
  if (u >= a.ngen_h * a.deg) return;
  u32 e = up_patch_low(a, u);
  if (e != INVALID) a.slot[u] = atomicAdd(&a.cnt[e], 1u);
}

// after the scan of the counts: groups of G lanes over the old lows write
// a surviving low's surviving uses in order (G at a time, placed by a
// prefix count over the group's ballot); the ngen_h * deg threads after
// them put each product use's index at its place
template <u32 G>
__global__ __launch_bounds__(BLOCK) void up_patch_fill_k(UpPatch a) {
  size_t gid = (size_t) xcd_block() * BLOCK + threadIdx.x;
  size_t const ngroups_lanes = (size_t) a.nl_old * G;
  if (gid < ngroups_lanes) {
    u32 const t = (u32) (gid / G), lane = (u32) (gid % G);
    u32 n = a.soff_l[t];
    if (n == a.soff_l[t + 1]) return;
    u32 const first = (u32) (threadIdx.x & 63u) & ~(G - 1u);  // the group's first lane in the wave
    unsigned long long const below = ((1ull << lane) - 1ull) << first;
    u32 p = a.off[n];
    for (u32 k0 = a.uo[t]; k0 < a.uo[t + 1]; k0 += G) {
      u32 const k = k0 + lane;
      bool keep = false;
      u32 s = 0;
      if (k < a.uo[t + 1]) {
        u32 h = a.ua[k];
        s = a.soff_h[h];
        keep = s != a.soff_h[h + 1];
      }
      unsigned long long const m = __ballot(keep);
      u32 const q = p + (u32) __popcll(m & below);
      if (keep && q < a.nuse) {
        a.adj[q] = s;
        a.dirs[q] = a.ud[k];
      }
      p += (u32) __popcll((m >> first) & ((G == 64 ? ~0ull : (1ull << G) - 1ull)));
    }
  } else if (gid - ngroups_lanes < (size_t) a.ngen_h * a.deg) {
    u32 u = (u32) (gid - ngroups_lanes);
    u32 e = up_patch_low(a, u);
    if (e == INVALID) return;
    u32 const p = a.off[e] + a.slot[u];
    if (p < a.nuse) a.key[p] = a.nsame_h * a.deg + u;
  }
}

// each product use's final place: its rank by use index among its low's
// product uses
__global__ __launch_bounds__(BLOCK) void up_patch_sort_k(UpPatch a) {
  u32 u = xcd_block() * BLOCK + threadIdx.x;
  if (u >= a.ngen_h * a.deg) return;
  u32 e = up_patch_low(a, u);
  if (e == INVALID) return;
  u32 const b = a.off[e] + a.cnt_s[e], end = a.off[e + 1] < a.nuse ? a.off[e + 1] : a.nuse;
  u32 const me = a.nsame_h * a.deg + u;
  u32 r = 0;
  for (u32 k = b; k < end; ++k) r += a.key[k] < me ? 1u : 0u;
  if (b + r >= a.nuse) return;
  u32 const h = me / a.deg;
  a.adj[b + r] = h;
  a.dirs[b + r] = me - a.deg * h;
}

// per new low, kept for the caller: all its uses (cnt), those in surviving
// highs (cnt_s), and whether those are all its old ones (full); a low with
// full set and cnt == cnt_s has exactly its old highs
struct UpPatchCounts {
  Buf<u32> cnt, cnt_s, full;
};

// the new mesh's up(l,h) from the old mesh's: nl_old old lows, survivors
// of both dimensions given by their offsets over the old entities, dn the
// new down(h,l) (its product rows are read)
static std::unique_ptr<Up> up_patch(u32 deg, Up const& old, u32 nl_old, u32 const* soff_l, u32 nsame_l,
                                    u32 ngen_l, u32 const* soff_h, u32 nsame_h, u32 ngen_h, u32 const* dn,
                                    UpPatchCounts* keep = nullptr) {
  u32 const nl = nsame_l + ngen_l, nuse = (nsame_h + ngen_h) * deg, nprod = ngen_h * deg;
  auto up = std::make_unique<Up>();
  up->offsets = Buf<u32>(nl + 1);
  up->adj = Buf<u32>(nuse);
  up->dirs = Buf<u32>(nuse);
  Buf<u32> cnt(nl), cnt_s(nl), slot(nprod), key(nuse), full;
  if (keep) full = Buf<u32>(nl);
  UpPatch a{};
  a.deg = deg;
  a.nuse = nuse;
  a.nl_old = nl_old;
  a.nsame_l = nsame_l;
  a.ngen_l = ngen_l;
  a.nsame_h = nsame_h;
  a.ngen_h = ngen_h;
  a.uo = old.offsets.get();
  a.ua = old.adj.get();
  a.ud = old.dirs.get();
  a.soff_l = soff_l;
  a.soff_h = soff_h;
  a.dn = dn;
  a.cnt = cnt.get();
  a.cnt_s = cnt_s.get();
  a.full = full.get();
  a.slot = slot.get();
  a.off = up->offsets.get();
  a.key = key.get();
  a.adj = up->adj.get();
  a.dirs = up->dirs.get();
  // compulsory bytes: the old lists and the highs' survivor flags read
  double const old_bytes = (double) nl_old * 12 + (double) old.adj.size() * 12;
  // lanes per old low: about its mean list length (edges in tets ~5,
  // vertices in edges ~14, vertices in tets ~24)
  size_t const mean = nl_old ? old.adj.size() / nl_old : 0;
  u32 const G = mean <= 6 ? 4u : mean <= 16 ? 8u : 16u;
  auto groups = [&](auto count_k, auto fill_k) {
    size_t const nc = ((size_t) nl_old + ngen_l) * G;
    if (nc) launch_timed("up_patch_count", old_bytes, count_k, dim3(nblocks(nc)), dim3(BLOCK), a);
    if (nprod) {
      hipLaunchKernelGGL(up_patch_slot_k, dim3(nblocks(nprod)), dim3(BLOCK), 0, stream(), a);
      check_launch("up_patch_slot");
    }
    exscan(cnt.get(), nl, up->offsets.get());
    size_t const nf = (size_t) nl_old * G + nprod;
    if (nf) launch_timed("up_patch_fill", old_bytes + (double) nuse * 8, fill_k, dim3(nblocks(nf)), dim3(BLOCK), a);
  };
  if (G == 4) groups(up_patch_count_k<4>, up_patch_fill_k<4>);
  else if (G == 8) groups(up_patch_count_k<8>, up_patch_fill_k<8>);
  else groups(up_patch_count_k<16>, up_patch_fill_k<16>);
  if (nprod) {
    hipLaunchKernelGGL(up_patch_sort_k, dim3(nblocks(nprod)), dim3(BLOCK), 0, stream(), a);
    check_launch("up_patch_sort");
  }
  if (keep) {
    keep->cnt = std::move(cnt);
    keep->cnt_s = std::move(cnt_s);
    keep->full = std::move(full);
  }
  return up;
}

static u32 coarsen_common(Mesh& m, Buf<u32>& codes, double qual_floor, bool require_better,
                          double const* qfe = nullptr, double const* eq_given = nullptr,
                          SizeCodeCache* sizes = nullptr, double size_floor = 0) {
  check_index_space(m);
  u32 dim = m.dim;
  u32 ne = count(m, 1);
  check_basic_class(m, codes.get());
  for (u32 d = 2; d <= dim; ++d) check_exposed_sides(m, d, codes.get());
  // the collapse directions left, listed vertex by vertex (each vertex's
  // edges in up(0,1) order): the quality kernel's groups then all have
  // work, and neighbouring groups share a vertex's star (the count, read
  // once, is also the early return when the class checks leave nothing).
  // The list is the uses of up(0,1) that collapse, in use order, so one
  // flat scan over the uses numbers them (no per-vertex loop), and a
  // vertex's first direction is the scan at its first use.
  u32 ncand;
  Buf<u32> clist, voff;
  u32 const nv0 = count(m, 0);
  {
    Up const& uv = ask_up(m, 0, 1);
    u32 const* uo = uv.offsets.get();
    u32 const* ua = uv.adj.get();
    u32 const* ud = uv.dirs.get();
    u32 const* cp = codes.get();
    u32 const nuse = 2 * ne;
    Buf<u32> uoff(nuse + 1);
    ncand = exscan_fn_total(nuse, OHX_LAMBDA(u32 k) { return collapses(cp[ua[k]], ud[k]); }, uoff.get());
    if (!ncand) return 0;
    clist = Buf<u32>(ncand);
    voff = Buf<u32>(nv0 + 1);
    u32 const* uf = uoff.get();
    u32* cl = clist.get();
    u32* vo = voff.get();
    parallel_for(nuse > nv0 ? nuse : nv0 + 1, OHX_LAMBDA(u32 k) {
      if (k < nuse && uf[k] != uf[k + 1]) cl[uf[k]] = 2 * ua[k] + ud[k];
      if (k <= nv0) vo[k] = uf[uo[k]];
    });
  }
  Buf<double> eq;
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vel = ask_down(m, dim, 0);
  Up const& ue = ask_up(m, 0, dim);
  // the element qualities of this mesh: the caller's (the adapt loop
  // computes them once per mesh) or evaluated here
  double const* eqp = nullptr;
  if (require_better) {
    if (!eq_given) eq = element_qualities(dim, count(m, dim), vel, coords(m), count(m, 0), metric_of(m));
    eqp = eq_given ? eq_given : eq.get();
  }
  Buf<double> cq((size_t) 2 * ne);
  Buf<u32> cand;
  Buf<double> vq;
  {
    double prof_bytes = (double) ne * 32 + (double) ue.adj.size() * 8 + (double) count(m, dim) * (dim + 1) * 4 + (double) count(m, 0) * 28;
    // variants, chosen online per size class: lanes per edge (1, 2, 4 or
    // 8) and, in the mixed mode, the FP32 screen (+16) or the certified FP32
    // pass (+32); all give identical results.  The mixed mode takes the
    // certified pass from CQ_CERTIFIED_MIN candidates up (below that its
    // three extra small launches cost more than the FP64 work it saves),
    // lanes per edge tuned; the screen stays available to pin.
    static thread_local Tuner tuners[2][3] = {
        {Tuner("coarsen_quality_2d_f64", {1, 2, 4, 8}),
         Tuner("coarsen_quality_2d_mixed", {33, 34, 36, 40, 17, 18, 20, 24}),
         Tuner("coarsen_quality_2d_mixed_small", {1, 2, 4, 8})},
        {Tuner("coarsen_quality_3d_f64", {1, 2, 4, 8}),
         Tuner("coarsen_quality_3d_mixed", {33, 34, 36, 40, 17, 18, 20, 24}),
         Tuner("coarsen_quality_3d_mixed_small", {1, 2, 4, 8})}};
    constexpr u32 CQ_CERTIFIED_MIN = 4096;
    bool const big = ncand >= CQ_CERTIFIED_MIN || tuners[dim == 3][1].forced();
    Tuner& tuner = tuners[dim == 3][g_mixed ? (big ? 1 : 2) : 0];
    CqArgs a{};
    a.ncand = ncand;
    a.list = clist.get();
    a.ev = ev;
    a.uo = ue.offsets.get();
    a.ua = ue.adj.get();
    a.ud = ue.dirs.get();
    a.vel = vel;
    a.x = coords(m);
    a.mt = metric_of(m);
    a.eq = eqp;
    a.qual_floor = qual_floor;
    a.qfe = qfe;
    a.codes = codes.get();
    a.out = cq.get();
    a.vmark = nullptr;
    a.st = screen_stats();
    // collapse_to_vert (collapses_to_verts.cpp:13-70) belongs to the timed
    // variant: the certified pass adds its certification, the FP64
    // measurement of what it could not certify and a second collapse_to_vert,
    // so a trial times the whole sequence (events recorded on the stream
    // around it) and the tuner weighs the variants by what they cost
    u32 const nv = count(m, 0);
    Up const& uv = ask_up(m, 0, 1);
    cand = Buf<u32>(nv);
    vq = Buf<double>(nv);
    tuned(tuner, ncand, [&](u32 var, hipEvent_t ea, hipEvent_t eb) {
      u32 g = var & 15u;
      u32 mode = (var & 32u) ? (u32) CQ_CERTIFIED : (var & 16u) ? (u32) CQ_SCREEN : (u32) CQ_F64;
      bool const certified = mode == CQ_CERTIFIED && !require_better;
      if (ea) OHX_CHECK(hipEventRecord(ea, stream()));
      dim == 3 ? cq_launch_mode<3>(mode, g, a, prof_bytes, nullptr, nullptr)
               : cq_launch_mode<2>(mode, g, a, prof_bytes, nullptr, nullptr);
      Buf<u32> vinex, vmark;
      if (certified) vinex = Buf<u32>(nv);
      auto to_vert = [&] {
        if (!nv) return;
        hipLaunchKernelGGL(collapse_to_vert_k, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, uv.offsets.get(),
                           uv.adj.get(), uv.dirs.get(), codes.get(), cq.get(), voff.get(), cand.get(), vq.get(),
                           certified ? vinex.get() : nullptr);
        check_launch("collapse_to_vert");
      };
      to_vert();
      if (certified && nv) {
        // the comparisons the FP32 goodness takes part in: the vertices
        // whose choices are not certified get exact values (their inexact
        // directions measured in FP64), then their best again
        vmark = Buf<u32>(nv);
        hipLaunchKernelGGL(coarsen_certify, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, uv.offsets.get(),
                           uv.adj.get(), uv.dirs.get(), codes.get(), cq.get(), ev, cand.get(), vq.get(), vinex.get(),
                           vmark.get());
        check_launch("coarsen_certify");
        CqArgs f = a;
        f.vmark = vmark.get();
        dim == 3 ? cq_launch_mode<3>(CQ_FORCE, g, f, prof_bytes, nullptr, nullptr)
                 : cq_launch_mode<2>(CQ_FORCE, g, f, prof_bytes, nullptr, nullptr);
        to_vert();
      }
      if (eb) OHX_CHECK(hipEventRecord(eb, stream()));
    });
  }
  // no host check here: if the quality check left no collapse, the
  // candidates below are empty, the independent set settles at once and the
  // one read of its total returns 0 (check_coarsen_quality's early return)
  u32 nv = count(m, 0);
  Up const& uv = ask_up(m, 0, 1);
  eq.release();
  // a first batch of indset rounds, the collapse destinations and the
  // vertex offsets, then one read of the convergence flag and the total
  u32 const* flag = nullptr;
  Buf<u32> indset = find_indset_bounded(m, 0, cand.get(), vq.get(), &flag);
  Buf<u32> gvv(nv);
  Buf<u32> goff(nv + 1);
  // coarsen_all_ents (coarsen_common.cpp:91-107), top dimension first.
  // Phase 1 decides every dimension's survivors and products on the device
  // (a dimension's fused sides feed the next lower one, no host decision
  // needed), right after the destinations; one read then brings the set's
  // convergence flag, the collapse count and all the totals; phase 2
  // writes the outputs.
  bool has_class = find_tag(m, 0, "class_dim") != nullptr;
  bool has_id = find_tag(m, 0, "class_id") != nullptr;
  Buf<u32> fused[4], fused_to[4];
  struct DimPlan {
    Buf<u32> gv, gd, egoff, esoff;
  };
  DimPlan plan[4];
  u32 got[10];
  auto dest_and_total = [&] {
    // vertex_dest (collapses_to_verts.cpp:72-135)
    u32 const* uo = uv.offsets.get();
    u32 const* ua = uv.adj.get();
    u32 const* ud = uv.dirs.get();
    u32 const* cd = codes.get();
    double const* q = cq.get();
    double const* vqp = vq.get();
    u32 const* is = indset.get();
    u32* g = gvv.get();
    u32* err = dev_error_word();
    u32 const* fl = flag;
    parallel_for(nv, OHX_LAMBDA(u32 i) {
      if (!is[i] || (fl && *fl)) {  // not in the set, or the set is not settled yet
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
      if (gv == INVALID) *err = ERR_NO_DESTINATION;
      g[i] = gv;
    });
    // coarsen_interior (coarsen_common.cpp:191-219)
    exscan(indset.get(), nv, goff.get());
    for (u32 dd = 0; dd < dim; ++dd) {
      u32 d = dim - dd;
      u32 n = count(m, d);
      u32 const* vent = ask_down(m, d, 0);
      u32 const* sent = ask_down(m, d, d - 1);
      u32 nsd = count(m, d - 1);
      fused[d - 1] = filled(nsd, 0);
      if (d >= 2) fused_to[d - 1] = Buf<u32>(nsd);
      u32 vpe = d + 1, spe = h_down_deg[d][d - 1];
      Buf<u32> will(n), gv(n), gd(n), same(n);
      {
        // collapse_to_ent (collapses_to_ents.cpp:11-82); also marks the
        // entities with no collapsing vertex, which survive unchanged, and
        // notes for each fused side the side it merges into (the one
        // opposite the collapsing vertex), which the maps carried to the
        // new mesh read
        u32 const* go = goff.get();
        u32 const* g = gvv.get();
        u32 const* fe = fused[d] ? fused[d].get() : nullptr;
        u32* fs = fused[d - 1].get();
        u32* ft = d >= 2 ? fused_to[d - 1].get() : nullptr;
        u32* wl = will.get();
        u32* gvp = gv.get();
        u32* gdp = gd.get();
        u32* smp = same.get();
        parallel_for(n, OHX_LAMBDA(u32 i) {
          u32 const* v = vent + (size_t) i * vpe;
          u32 col = INVALID, dir = INVALID;
          for (u32 j = 0; j < vpe; ++j)
            if (go[v[j]] != go[v[j] + 1]) {
              col = v[j];
              dir = j;
            }
          u32 w = 0, gvv_ = 0, gd_ = 0;
          if (dir != INVALID) {
            u32 target = g[col];
            u32 gdir = INVALID;
            for (u32 j = 0; j < vpe; ++j)
              if (v[j] == target) gdir = j;
            if (gdir != INVALID) {
              u32 const side = sent[(size_t) i * spe + opp(d, 0)[gdir]];
              fs[side] = 1;
              if (ft) ft[side] = sent[(size_t) i * spe + opp(d, 0)[dir]];
            } else if (!(fe && fe[i])) {
              w = 1;
              gvv_ = target;
              gd_ = dir;
            }
          }
          wl[i] = w;
          gvp[i] = gvv_;
          gdp[i] = gd_;
          smp[i] = dir == INVALID ? 1u : 0u;
        });
      }
      Buf<u32> egoff(n + 1), esoff(n + 1);
      {
        u32 const* wl = will.get();
        u32 const* smp = same.get();
        exscan_fn(n, OHX_LAMBDA(u32 i) { return wl[i]; }, egoff.get());
        exscan_fn(n, OHX_LAMBDA(u32 i) { return smp[i]; }, esoff.get());
      }
      same.release();
      plan[d] = DimPlan{std::move(gv), std::move(gd), std::move(egoff), std::move(esoff)};
    }
    u32 const* ptrs[10];
    u32 k = 0;
    ptrs[k++] = flag;
    ptrs[k++] = goff.get() + nv;
    for (u32 d = 1; d <= dim; ++d) {
      u32 n = count(m, d);
      ptrs[k++] = plan[d].egoff.get() + n;
      ptrs[k++] = plan[d].esoff.get() + n;
    }
    read_many(ptrs, k, got);
  };
  dest_and_total();
  if (got[0]) {  // not settled in the first batch: the checked loop from there
    indset = find_indset_resume(m, 0, vq.get(), std::move(indset));
    flag = nullptr;
    dest_and_total();
  }
  if (g_batch) batch_note_changed(m, 0, indset.get());
  cand.release();
  codes.release();
  cq.release();
  vq.release();
  u32 total = got[1];
  if (!total) return 0;
  indset.release();
  Buf<u32> soff = negate_offsets(goff.get(), nv);
  u32 nv_out = nv - total;
  Mesh mo(dim, m.full);
  set_ents(mo, 0, nv_out, Buf<u32>());
  {
    // tags_subset (subset.cpp:15-43): every vertex tag, one launch
    TagList L;
    for (auto& tp : m.tags[0]) {
      Tag& t = *tp;
      void* dst;
      if (t.type == TagType::F64)
        dst = add_tag_f64(mo, 0, t.name.c_str(), t.ncomps, Buf<double>((size_t) nv_out * t.ncomps), t.transfer).f.get();
      else
        dst = add_tag_u32(mo, 0, t.name.c_str(), t.ncomps, Buf<u32>((size_t) nv_out * t.ncomps), t.transfer).u.get();
      tag_list_add(L, t, dst);
    }
    if (L.n) {
      hipLaunchKernelGGL(subset_tags, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, soff.get(), L);
      check_launch("subset_tags");
    }
  }
  // the maps this pass read and the next collapse pass reads again,
  // carried to the new mesh instead of derived again: each entity's sides
  // (down(d, d-1)) here, the vertices' upward adjacencies below.  The tets'
  // edges and up(1,3) are carried by the sliver collapse passes only: a swap
  // pass reads them, and carrying them through the collapse-by-size loop
  // as well cost more than one reflection before the swap (64^3 x 6
  // anisotropic box: 2.87 against 2.82 s per 8 cycles)
  Buf<u32> side_maps[4][4];
  for (u32 dd = 0; dd < dim; ++dd) {
    u32 d = dim - dd;
    u32 n = count(m, d);
    u32 vpe = d + 1;
    u32 ngen = got[2 + 2 * (d - 1)], nsame = got[2 + 2 * (d - 1) + 1];
    DimPlan& pl = plan[d];
    Buf<u32> vout((size_t) (nsame + ngen) * vpe);
    Buf<u32> cdo, cio;
    CoarsenArgs a{};
    auto carry_sides = [&](u32 l, CoarsenSides& cs) {
      u32 const ns = simplex_sides(d, l);
      side_maps[d][l] = Buf<u32>((size_t) (nsame + ngen) * ns);
      cs.out = side_maps[d][l].get();
      cs.old = m.down[d][l].get();
      cs.soff = plan[l].esoff.get();
      cs.egoff = plan[l].egoff.get();
      cs.fused_to = fused_to[l].get();
      cs.nsame = got[2 + 2 * (l - 1) + 1];
      cs.l = l;
      cs.nold = m.counts[l];
      cs.err = dev_error_word();
    };
    if (d >= 2 && m.down[d][d - 1]) carry_sides(d - 1, a.sides);
    // a sliver collapse pass is followed by a swap pass, which reads the
    // tets' edges (and up(1,3), carried below)
    if (require_better && d == 3 && m.down[3][1]) carry_sides(1, a.edges);
    a.d = d;
    a.n = n;
    a.nsame = nsame;
    a.vent = ask_down(m, d, 0);
    a.esoff = pl.esoff.get();
    a.egoff = pl.egoff.get();
    a.gv = pl.gv.get();
    a.gd = pl.gd.get();
    a.vsoff = soff.get();
    a.out = vout.get();
    if (has_class) {
      cdo = Buf<u32>(nsame + ngen);
      a.cd = u32_tag(m, d, "class_dim");
      a.out_cd = cdo.get();
      if (has_id) {
        cio = Buf<u32>(nsame + ngen);
        a.ci = u32_tag(m, d, "class_id");
        a.out_ci = cio.get();
      }
    }
    // compulsory bytes per old entity: its vertices in, their new numbers
    // gathered and written, its survivor and product offsets, target and
    // direction, class tags in and out, and the carried sides (old ids in,
    // new ids gathered and written)
    u32 const ns = (a.sides.out ? simplex_sides(d, d - 1) : 0u) + (a.edges.out ? 6u : 0u);
    launch_timed("coarsen_outputs", (double) n * (vpe * 12 + 32 + ns * 12), coarsen_outputs, dim3(nblocks(n)),
                 dim3(BLOCK), a);
    set_ents(mo, d, nsame + ngen, std::move(vout));
    if (has_class) {
      add_tag_u32(mo, d, "class_dim", 1, std::move(cdo));
      if (has_id) add_tag_u32(mo, d, "class_id", 1, std::move(cio));
    }
    if (d == dim) coarsen_transfer(m, mo, goff.get(), total, pl.egoff.get(), pl.esoff.get(), nsame, ngen);
  }
  for (u32 h = 2; h <= dim; ++h)
    for (u32 l = 1; l < h; ++l)
      if (side_maps[h][l]) mo.down[h][l] = std::move(side_maps[h][l]);
  // the vertices' upward adjacencies this pass read (to edges, to
  // elements), carried too (vertices survive or are removed, none is new)
  auto carry_up = [&](u32 h) {
    if (!m.up[0][h]) return;
    mo.up[0][h] = up_patch(h_down_deg[h][0], *m.up[0][h], m.counts[0], soff.get(), nv_out, 0, plan[h].esoff.get(),
                           got[2 + 2 * (h - 1) + 1], got[2 + 2 * (h - 1)], mo.down[h][0].get());
  };
  carry_up(1);
  carry_up(dim);
  if (sizes && sizes->ne == m.counts[1]) {
    // the new mesh's size codes: survivors' carried, products measured
    u32 const ne_old = m.counts[1], ngen1 = got[2], nsame1 = got[3];
    Buf<u32> nc((size_t) nsame1 + ngen1);
    hipLaunchKernelGGL(carry_size_codes_k, dim3(nblocks((size_t) ne_old + ngen1)), dim3(BLOCK), 0, stream(), ne_old,
                       ngen1, nsame1, plan[1].esoff.get(), sizes->codes.get(), mo.down[1][0].get(), coords(mo),
                       size_of(mo), metric_of(mo), size_floor, nc.get());
    check_launch("carry_size_codes");
    sizes->ne = nsame1 + ngen1;
    sizes->codes = std::move(nc);
  } else if (sizes) {
    sizes->clear();
  }
  if (require_better && dim == 3 && m.up[1][3] && mo.down[3][1])
    mo.up[1][3] = up_patch(6, *m.up[1][3], m.counts[1], plan[1].esoff.get(), got[3], got[2], plan[3].esoff.get(),
                           got[7], got[6], mo.down[3][1].get());
  say("collapsed", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 coarsen_by_size(Mesh& m, double qual_floor, double size_floor, SizeCodeCache* cache) {
  u32 ne = count(m, 1);
  double const* h = size_of(m);
  double const* mt = metric_of(m);
  Buf<u32> codes(ne);
  u32 tag = next_any_tag();
  if (g_batch && cache) cache->clear();
  if (cache && cache->ne != ne) cache->clear();
  if (ne && cache && cache->ne) {
    // the codes carried from the pass before (the pass consumes its copy)
    u32* c = codes.get();
    u32 const* cc = cache->codes.get();
    u32* words = scratch_words();
    parallel_for(ne, OHX_LAMBDA(u32 i) {
      u32 const v = cc[i];
      c[i] = v;
      wave_flag_value(words + W_ANY, v != 0, tag);
    });
  } else if (ne) {
    double prof_bytes = (double) ne * 12 + (double) count(m, 0) * 32;
    launch_timed("coarsen_size_code_f64", prof_bytes, coarsen_size_code_f64, dim3(nblocks(ne)), dim3(BLOCK), ne,
                       ask_down(m, 1, 0), coords(m), h, mt, size_floor, codes.get(), scratch_words(), tag);
    if (cache && !g_batch) {
      cache->ne = ne;
      cache->codes = copy(codes.get(), ne);
    }
  }
  // check_coarsen_noop (coarsen_common.cpp:109-118), kept as its own read:
  // without it the last, unproductive pass of every coarsening loop would
  // derive the maps of the class checks for nothing, which on small meshes
  // costs more than the host wait it saves on each productive pass
  if (read_word(W_ANY) != tag) return 0;
  if (g_batch) {
    batch_mask(m, 1, codes.get());
    Buf<double> qfe = batch_floors(m, 1);
    return coarsen_common(m, codes, qual_floor, false, qfe.get());
  }
  return coarsen_common(m, codes, qual_floor, false, nullptr, nullptr, cache, size_floor);
}

// coarsen_slivers (coarsen.cpp:38-70)
u32 coarsen_slivers(Mesh& m, double qual_floor, u32 nlayers, double const* eq) {
  u32 dim = m.dim;
  Buf<u32> sl = mark_slivers(m, qual_floor, nlayers, eq);
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
  if (g_batch) batch_mask(m, 1, c);
  // check_coarsen_noop (coarsen_common.cpp:109-118) is coarsen_common's
  // count of the directions left (its first host read)
  return coarsen_common(m, codes, 0.0, true, nullptr, eq);
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
static std::mutex g_swap_mu;

// The interior edges of each triangulation, derived the way the
// reference's generator does (test_print_swap_edges.cpp:29-64): the
// triangulation as a reduced 2D mesh, its derived edges, the non-boundary
// ones in edge order.  All triangulations of one N form one mesh of
// disjoint components; side derivation numbers each component's edges
// contiguously and in the same order as a mesh of its own would.
static void make_swap_tables() {
  std::lock_guard<std::mutex> lock(g_swap_mu);
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

// swap_quality (swap_qualities.cpp:13-65) with choose_edge_swap
// (edge_swap.cpp:230-280), SW_G lanes per edge.  One lane per edge would
// hold the ring's coordinates and metrics and the quality of every unique
// triangle (up to 35) in registers, which spills; here a group of lanes
// shares them in LDS:
//   A. lane j < ring size reads the edge's j-th tet: its opposite edge
//      (a, b) oriented as find_edge_ring orients it, and its quality
//      (old_minq is the minimum over the group);
//   B. lane 0 chains the pairs into the ring (edge_ring.cpp:12-74);
//   C. lane j stages ring vertex j (or edge vertex j - ring size): its
//      coordinates and metric;
//   D. lane l measures unique triangles l, l + SW_G, ...: the two tets it
//      forms with the edge's ends, as choose_edge_swap does;
//   E. lane l takes triangulations l, l + SW_G, ...: the running minimum of
//      its triangles, stopped at the first negative one as the reference's
//      loop stops, and the lane's best with a strict > in triangulation
//      order; the group combines them keeping the higher value and, on
//      ties, the lower triangulation index -- the reference's sequential
//      strict > choice.
// FP32-mixed (MIXED): D and E run first with the FP32 quality; a candidate
// whose FP32 best lies more than the margin below old_minq (and none of
// whose triangles is within the margin of zero, where the reference's
// early stop could fall elsewhere) is rejected there, every other one is
// measured again in FP64, so every accepted value and code is the FP64
// path's.  Accepted candidates store the value (goodness), the ring size
// and the triangulation code, which swap_cavity uses.  The kernel is built
// for 4 waves per SIMD (at most 128 VGPRs; its 33 KB of LDS per block allow
// 4 blocks per CU): with a metric that costs 28 bytes of scratch per lane
// and measured 4 % less time on the 64^3 x 6 anisotropic box than 3 waves without spills.
constexpr u32 SW_G = 8;
constexpr u32 SW_EDGES = BLOCK / SW_G;
constexpr u32 SW_MAXTRI = 35;
static __constant__ u32 k_sw_ntri[8] = {0, 0, 0, 1, 4, 10, 20, 35};

template <bool MT>
struct SwapLds {
  u32 ra[SW_EDGES][MAX_EDGE_SWAP], rb[SW_EDGES][MAX_EDGE_SWAP];
  u32 ring[SW_EDGES][MAX_EDGE_SWAP];
  u32 rs[SW_EDGES];      // ring size of an active edge, else 0
  u32 again[SW_EDGES];   // FP32-mixed: measure again in FP64
  double old_minq[SW_EDGES];
  double x[SW_EDGES][MAX_EDGE_SWAP + 2][3];          // ring, then edge vertices
  double m[MT ? SW_EDGES : 1][MAX_EDGE_SWAP + 2][6];  // their metrics
  double tq[SW_EDGES][SW_MAXTRI];
};

template <bool MT, bool F32>
__device__ inline double swap_tet_q(SwapLds<MT> const& L, u32 g, u32 a0, u32 a1, u32 a2, u32 e) {
  V3 p[4];
  u32 const idx[4] = {a0, a1, a2, e};
#pragma unroll
  for (u32 k = 0; k < 4; ++k) p[k] = V3{L.x[g][idx[k]][0], L.x[g][idx[k]][1], L.x[g][idx[k]][2]};
  if (MT) {
    Sym pm[4];
#pragma unroll
    for (u32 k = 0; k < 4; ++k)
#pragma unroll
      for (u32 c = 0; c < 6; ++c) pm[k].m[c] = L.m[g][idx[k]][c];
    return F32 ? (double) quality32<3>(p, pm, true) : metric_quality<3>(p, pm, true);
  }
  return F32 ? (double) quality32<3>(p, nullptr, true) : tet_quality(p[0], p[1], p[2], p[3]);
}

// phases D and E for the group of this lane; returns the group's (best,
// code) in every lane of the group, *near0 set if any triangle's quality is
// within SCREEN_MARGIN of zero
template <bool MT, bool F32>
__device__ inline void swap_choose(SwapLds<MT>& L, u32 g, u32 lane, bool active, double* best, u32* code,
                                   bool* near0) {
  u32 rs = active ? L.rs[g] : 0u;
  bool nz = false;
  if (rs) {
    u32 ntri = k_sw_ntri[rs];
    unsigned char const(*tris)[3] = k_sw_tris + k_sw_tri_off[rs];
    u32 const E0 = rs, E1 = rs + 1;  // edge vertices follow the ring in LDS
    for (u32 t = lane; t < ntri; t += SW_G) {
      u32 a0 = tris[t][0], a1 = tris[t][1], a2 = tris[t][2];
      double tmin = 1, q;
      // the triangle with edge end 1, then with the first two corners
      // exchanged and edge end 0 (choose_edge_swap's two tets)
      q = swap_tet_q<MT, F32>(L, g, a0, a1, a2, E1);
      if (q < tmin) tmin = q;
      q = swap_tet_q<MT, F32>(L, g, a1, a0, a2, E0);
      if (q < tmin) tmin = q;
      if (F32 && !(tmin > (double) SCREEN_MARGIN || tmin < -(double) SCREEN_MARGIN)) nz = true;
      L.tq[g][t] = tmin;
    }
  }
  __syncthreads();
  double b = -1;
  u32 c = INVALID;
  if (rs) {
    u32 tpm = k_sw_sizes[rs], nm = k_sw_counts[rs];
    unsigned char const* meshes = k_sw_meshes + k_sw_mesh_off[rs];
    for (u32 k = lane; k < nm; k += SW_G) {
      double mq = 1;
      for (u32 j = 0; j < tpm; ++j) {
        double tq = L.tq[g][meshes[k * tpm + j]];
        if (tq < mq) mq = tq;
        if (mq < 0) break;
      }
      if (mq > b) {
        b = mq;
        c = k;
      }
    }
  }
  // combine over the group: higher value, then lower triangulation index
#pragma unroll
  for (u32 d = SW_G / 2; d >= 1; d >>= 1) {
    double ob = __shfl_xor(b, d, SW_G);
    u32 oc = (u32) __shfl_xor((int) c, d, SW_G);
    if (ob > b || (ob == b && oc < c)) {
      b = ob;
      c = oc;
    }
    nz = nz || __shfl_xor((int) nz, d, SW_G);
  }
  *best = b;
  *code = c;
  *near0 = nz;
}

template <bool MT, bool MIXED>
__global__ __launch_bounds__(BLOCK, 4) void swap_quality_k(
    u32 ncand, u32 const* __restrict__ list, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
    u32 const* __restrict__ ud, u32 const* __restrict__ ev, u32 const* __restrict__ vt,
    double const* __restrict__ x, double const* __restrict__ mt, double const* __restrict__ eq,
    u32* __restrict__ cand, double* __restrict__ out, u32* __restrict__ rsz, u32* __restrict__ codes,
    u32* __restrict__ words, u32 tag, u32* __restrict__ err, ScreenStats st) {
  __shared__ SwapLds<MT> L;
  u32 const g = threadIdx.x / SW_G, lane = threadIdx.x % SW_G;
  // the candidates, compacted: group g of the block takes one
  u32 const slot = xcd_block() * SW_EDGES + g;
  bool active = slot < ncand;
  u32 const i = active ? list[slot] : 0u;
  // A: the tets around the edge
  u32 f = 0, rs = 0, e0 = 0, e1 = 0;
  double qmin = 1;
  if (active) {
    f = uo[i];
    rs = uo[i + 1] - f;
    e0 = ev[2 * (size_t) i];
    e1 = ev[2 * (size_t) i + 1];
    if (rs > MAX_EDGE_SWAP || rs < 3) {
      if (lane == 0) {
        if (rs < 3) *err = ERR_EDGE_RING;
        cand[i] = 0;
      }
      active = false;
    } else if (lane < rs) {
      u32 tet = ua[f + lane], te = ud[f + lane];
      u32 const* tv = vt + 4 * (size_t) tet;
      u32 const* tve = canon(3, 1, 0, te);
      u32 al = (tv[tve[0]] != e0) ? 1u : 0u;
      u32 const* tvo = canon(3, 1, 0, opp(3, 1)[te]);
      L.ra[g][lane] = tv[tvo[al]];
      L.rb[g][lane] = tv[tvo[al ^ 1u]];
      double q = eq[tet];
      if (q < qmin) qmin = q;
    }
  }
#pragma unroll
  for (u32 d = SW_G / 2; d >= 1; d >>= 1) {
    double o = __shfl_xor(qmin, d, SW_G);
    if (o < qmin) qmin = o;
  }
  __syncthreads();
  // B: the ring, chained as find_edge_ring does
  if (active && lane == 0) {
    u32* ra = L.ra[g];
    u32* rb = L.rb[g];
    bool closed = true;
    for (u32 a = 0; a + 1 < rs; ++a) {
      u32 j;
      for (j = a + 1; j < rs; ++j)
        if (ra[j] == rb[a]) break;
      if (j == rs) {
        closed = false;
        break;
      }
      u32 ta = ra[j], tb = rb[j];
      ra[j] = ra[a + 1];
      rb[j] = rb[a + 1];
      ra[a + 1] = ta;
      rb[a + 1] = tb;
    }
    if (!closed) {
      *err = ERR_EDGE_RING;
      cand[i] = 0;
    }
    L.rs[g] = closed ? rs : 0u;
    L.old_minq[g] = qmin;
    for (u32 a = 0; a < rs; ++a) L.ring[g][a] = ra[a];
  } else if (lane == 0) {
    L.rs[g] = 0;
  }
  __syncthreads();
  // C: the ring's and the edge's vertices
  rs = active ? L.rs[g] : 0u;
  active = rs != 0;
  // (a 7-ring has 9 vertices: more than the group's lanes)
  for (u32 k = lane; active && k < rs + 2; k += SW_G) {
    u32 v = k < rs ? L.ring[g][k] : (k == rs ? e0 : e1);
#pragma unroll
    for (u32 c = 0; c < 3; ++c) L.x[g][k][c] = x[3 * (size_t) v + c];
    if (MT)
#pragma unroll
      for (u32 c = 0; c < 6; ++c) L.m[g][k][c] = mt[6 * (size_t) v + c];
  }
  if (lane == 0) L.again[g] = 0;
  // the block's work needs the staged data and every lane of the group
  // (and of the block: the phases below are separated by barriers)
  __syncthreads();
  double best;
  u32 code;
  bool near0;
  double const old_minq = active ? L.old_minq[g] : 1.0;
  if (MIXED) {
    swap_choose<MT, true>(L, g, lane, active, &best, &code, &near0);
    bool reject = active && !near0 && best + (double) SCREEN_MARGIN <= old_minq;
    if (lane == 0) {
      if (active) L.again[g] = reject ? 0u : 1u;
    }
    wave_count(st.screened, lane == 0 && reject);
    wave_count(st.reeval, lane == 0 && active && !reject);
    if (reject && lane == 0) cand[i] = 0;
    __syncthreads();
    double best32 = best;
    bool const valid32 = !near0;  // no triangle near zero: best32 is the FP64 choice's, rounded
    bool again = active && L.again[g] != 0;
    swap_choose<MT, false>(L, g, lane, again, &best, &code, &near0);
    if (again && valid32 && lane == 0 && best == best) {
      double dd = best - best32;
      note_max_err(st.max_err, (float) (dd < 0 ? -dd : dd));
    }
    active = again;
  } else {
    swap_choose<MT, false>(L, g, lane, active, &best, &code, &near0);
  }
  bool accept = active && best > old_minq;
  if (active && lane == 0) {
    if (accept) {
      out[i] = best;
      rsz[i] = rs;
      codes[i] = code;
    } else {
      cand[i] = 0;
    }
  }
  wave_flag_value(words + W_ANY, accept, tag);
}

// count_swap_ents (edge_swap.cpp:375-386)
static inline __host__ __device__ u32 count_swap_ents(u32 rs, u32 d, u32 nint, u32 tpm) {
  return d == 1 ? nint : (d == 2 ? tpm + 2 * nint : 2 * tpm);
}

// The swapped mesh's tet-to-edge map, built from the old one instead of by
// reflection (reflect_down.cpp:13-81 gives the same map: an entity's lows
// are unique in a conforming mesh).  A surviving tet is adjacent to no
// swapped edge, so its six edges survive too and it keeps them, renumbered.
// A product tet's edges are its cavity's: ring edges and spokes (ring
// vertex to an end of the swapped edge), both edges of the old tets around
// the swapped edge, and the new interior edges.  out null: not built.
struct SwapEdgesOf {
  u32* out;            // (surviving + product tets) x 6 new edge ids
  u32 const* old;      // the old tets' edges
  u32 const* soff1;    // old edge -> its index among the surviving edges
  u32 const* goff1;    // interior edges generated per swapped edge
  u32 nsame1, nsame3;  // surviving edges, surviving tets
  u32* err;
  // the element dual carried too (dual null: not built): the old dual,
  // the tets' survivor offsets, per old tet face the product now across it
  // (written for the cavities' outer faces), and per product face whose
  // outside tet is in another cavity the old (tet, face) to resolve
  u32* dual;
  u32 const* old_dual;
  u32 const* soff3;
  u32* face_prod;
  u32* deferred;
};

constexpr u32 DUAL_PENDING = 0xfffffffeu;

// local face of a tet (vertices tv) with the vertices a, b, c
__device__ inline u32 tet_face_of(u32 const* tv, u32 a, u32 b, u32 c) {
  for (u32 f = 0; f < 4; ++f) {
    u32 const* fv = canon(3, 2, 0, f);
    u32 hits = 0;
    for (u32 q = 0; q < 3; ++q) {
      u32 v = tv[fv[q]];
      hits += (v == a || v == b || v == c) ? 1u : 0u;
    }
    if (hits == 3) return f;
  }
  return 0;
}

// local edge of a tet between local vertices a and b
__device__ inline u32 tet_edge_between(u32 a, u32 b) {
  for (u32 j = 0; j < 6; ++j) {
    u32 const* c = canon(3, 1, 0, j);
    if ((c[0] == a && c[1] == b) || (c[0] == b && c[1] == a)) return j;
  }
  return 0;
}

// swap_cavity (swap_topology.cpp:46-75) with get_swap_ents
// (edge_swap.cpp:282-373), every dimension in one launch: the ring again
// (integer work) and the triangulation code that swap_quality chose, then
// the products' vertices (interior edges, their faces, the tets: triangle t
// of the triangulation with edge end 1, then with its first two corners
// exchanged and edge end 0), written where out_d is given.  With E.out, the
// product tets' edges too: a product tet's vertices are labelled by where
// they come from (ring slot, or edge end); an edge between adjacent ring
// slots is a ring edge, one between a slot and an edge end a spoke, both
// edges of the old tet at that ring slot (the one holding ring vertices k
// and k + 1), and any other is one of the new interior edges.  The cavity's
// ring edges and spokes are staged in LDS per thread (indexed by slot).
__global__ __launch_bounds__(BLOCK) void swap_products_k(
    u32 ne, u32 const* __restrict__ indset, u32 const* __restrict__ goff1, u32 const* __restrict__ goff2,
    u32 const* __restrict__ goff3, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
    u32 const* __restrict__ ud, u32 const* __restrict__ ev, u32 const* __restrict__ vt,
    u32 const* __restrict__ codes, u32* __restrict__ out1, u32* __restrict__ out2, u32* __restrict__ out3,
    SwapEdgesOf E) {
  __shared__ u32 s_ring[MAX_EDGE_SWAP][BLOCK];
  __shared__ u32 s_spoke[2][MAX_EDGE_SWAP][BLOCK];
  __shared__ u32 s_tet[MAX_EDGE_SWAP][BLOCK];
  u32 const tid = threadIdx.x;
  u32 i = xcd_block() * BLOCK + tid;
  if (i >= ne || !indset[i]) return;  // no barrier follows
  u32 edge_v[2], ring_v[MAX_EDGE_SWAP];
  u32 rs = edge_ring(i, uo, ua, ud, ev, vt, edge_v, ring_v);
  if (rs < 3 || rs > MAX_EDGE_SWAP) return;  // cannot happen for an indset edge
  u32 const code = codes[i];
  u32 const tpm = k_sw_sizes[rs];
  unsigned char const* mesh = k_sw_meshes + k_sw_mesh_off[rs] + code * tpm;
  unsigned char const(*tris)[3] = k_sw_tris + k_sw_tri_off[rs];
  u32 const nint = k_sw_nint[rs];
  unsigned char const* ie = k_sw_int + k_sw_int_off[rs] + code * nint * 2;
  if (out1) {
    u32* o = out1 + (size_t) goff1[i] * 2;
    for (u32 k = 0; k < nint; ++k) {
      o[2 * k] = ring_v[ie[2 * k]];
      o[2 * k + 1] = ring_v[ie[2 * k + 1]];
    }
  }
  if (out2) {
    u32* o = out2 + (size_t) goff2[i] * 3;
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
  u32 const f = uo[i], e = uo[i + 1];
  for (u32 k = 0; E.out && k < rs; ++k) {
    u32 const a = ring_v[k], b = ring_v[k + 1 < rs ? k + 1 : 0];
    u32 tet = INVALID;
    for (u32 q = f; q < e && tet == INVALID; ++q) {
      u32 const* tv = vt + 4 * (size_t) ua[q];
      bool ha = false, hb = false;
      for (u32 c = 0; c < 4; ++c) {
        ha = ha || tv[c] == a;
        hb = hb || tv[c] == b;
      }
      if (ha && hb) tet = ua[q];
    }
    if (tet == INVALID) {
      *E.err = ERR_REFLECT_NOT_FOUND;
      return;
    }
    u32 const* tv = vt + 4 * (size_t) tet;
    u32 const* te = E.old + 6 * (size_t) tet;
    u32 la = 0, lb = 0, l0 = 0, l1 = 0;
    for (u32 c = 0; c < 4; ++c) {
      if (tv[c] == a) la = c;
      if (tv[c] == b) lb = c;
      if (tv[c] == edge_v[0]) l0 = c;
      if (tv[c] == edge_v[1]) l1 = c;
    }
    s_tet[k][tid] = tet;
    s_ring[k][tid] = E.soff1[te[tet_edge_between(la, lb)]];
    s_spoke[0][k][tid] = E.soff1[te[tet_edge_between(la, l0)]];
    s_spoke[1][k][tid] = E.soff1[te[tet_edge_between(la, l1)]];
  }
  if (!out3) return;
  u32* o3 = out3 + (size_t) goff3[i] * 4;
  u32 const first = E.nsame3 + goff3[i];
  for (u32 t = 0; t < tpm; ++t) {
    u32 lab[4], tv[4];  // labels: ring slot, or 8 + edge end; vertices
    for (u32 j = 0; j < 3; ++j) {
      lab[j] = tris[mesh[t]][j];
      tv[j] = ring_v[lab[j]];
    }
    for (u32 j = 0; j < 2; ++j) {
      lab[3] = 8 + (1 - j);
      tv[3] = edge_v[1 - j];
      for (u32 k = 0; k < 4; ++k) *o3++ = tv[k];
      if (!E.out) {
        u32 tmp = tv[0];
        tv[0] = tv[1];
        tv[1] = tmp;
        tmp = lab[0];
        lab[0] = lab[1];
        lab[1] = tmp;
        continue;
      }
      u32* eo = E.out + 6 * (size_t) (first + 2 * t + j);
      for (u32 le = 0; le < 6; ++le) {
        u32 const* c = canon(3, 1, 0, le);
        u32 a = lab[c[0]], b = lab[c[1]];
        u32 id = INVALID;
        if (a >= 8 || b >= 8) {
          if (a < 8) id = s_spoke[b - 8][a][tid];
          else if (b < 8) id = s_spoke[a - 8][b][tid];
        } else if (a + 1 == b || a == b + rs - 1) {
          id = s_ring[a][tid];
        } else if (b + 1 == a || b == a + rs - 1) {
          id = s_ring[b][tid];
        } else {
          for (u32 k = 0; k < nint; ++k)
            if ((ie[2 * k] == a && ie[2 * k + 1] == b) || (ie[2 * k] == b && ie[2 * k + 1] == a))
              id = E.nsame1 + E.goff1[i] + k;
        }
        if (id == INVALID) *E.err = ERR_REFLECT_NOT_FOUND;
        eo[le] = id;
      }
      if (E.dual) {
        // the tet across each face: the twin (same triangle, other end)
        // across the triangle; across a face on a ring edge, the tet
        // outside the ring tet's face there; across a face on an interior
        // edge, the product of the other triangle holding that edge
        u32 const self = first + 2 * t + j;
        u32* dd = E.dual + 4 * (size_t) self;
        for (u32 f = 0; f < 4; ++f) {
          u32 const* fv = canon(3, 2, 0, f);
          u32 r[2], nr = 0;
          bool has_end = false;
          for (u32 q = 0; q < 3; ++q) {
            u32 l = lab[fv[q]];
            if (l >= 8)
              has_end = true;
            else if (nr < 2)
              r[nr++] = l;
          }
          u32 nb = INVALID;
          if (!has_end) {
            nb = first + 2 * t + (1 - j);
          } else {
            u32 const a = r[0], b = r[1];
            bool const ring = a + 1 == b || a == b + rs - 1 || b + 1 == a || b == a + rs - 1;
            if (ring) {
              u32 const k = (a + 1 == b || a == b + rs - 1) ? a : b;
              u32 const tk = s_tet[k][tid];
              u32 const x = 1 - j;  // this tet's edge end
              u32 const lf = tet_face_of(vt + 4 * (size_t) tk, edge_v[x], ring_v[k], ring_v[k + 1 < rs ? k + 1 : 0]);
              E.face_prod[4 * (size_t) tk + lf] = self;
              u32 const o = E.old_dual[4 * (size_t) tk + lf];
              if (o == INVALID) {
                nb = INVALID;
              } else if (E.soff3[o] != E.soff3[o + 1]) {
                nb = E.soff3[o];
              } else {
                nb = DUAL_PENDING;
                E.deferred[4 * (size_t) (self - E.nsame3) + f] = 4 * tk + lf;
              }
            } else {
              for (u32 t2 = 0; t2 < tpm; ++t2) {
                if (t2 == t) continue;
                unsigned char const* tr = tris[mesh[t2]];
                bool ha = tr[0] == a || tr[1] == a || tr[2] == a;
                bool hb = tr[0] == b || tr[1] == b || tr[2] == b;
                if (ha && hb) nb = first + 2 * t2 + j;
              }
              if (nb == INVALID) *E.err = ERR_REFLECT_NOT_FOUND;
            }
          }
          dd[f] = nb;
        }
      }
      u32 tmp = tv[0];
      tv[0] = tv[1];
      tv[1] = tmp;
      tmp = lab[0];
      lab[0] = lab[1];
      lab[1] = tmp;
    }
  }
}

// The new dual's remaining entries: threads [0, nt_old) give a surviving
// tet its neighbours (a surviving one renumbered, a removed one replaced by
// the product now across that face); the ngen3 * 4 after them resolve a
// product face whose outside tet was removed by another cavity.
__global__ __launch_bounds__(BLOCK) void swap_dual_k(u32 nt_old, u32 ngen3, SwapEdgesOf E) {
  u32 const s = xcd_block() * BLOCK + threadIdx.x;
  auto across = [&](u32 n, u32 from) -> u32 {  // the product across n's face shared with from
    for (u32 g = 0; g < 4; ++g)
      if (E.old_dual[4 * (size_t) n + g] == from) return E.face_prod[4 * (size_t) n + g];
    *E.err = ERR_REFLECT_NOT_FOUND;
    return INVALID;
  };
  if (s < nt_old) {
    u32 const o = E.soff3[s];
    if (o == E.soff3[s + 1]) return;
    for (u32 f = 0; f < 4; ++f) {
      u32 const n = E.old_dual[4 * (size_t) s + f];
      u32 r = INVALID;
      if (n != INVALID) r = E.soff3[n] != E.soff3[n + 1] ? E.soff3[n] : across(n, s);
      E.dual[4 * (size_t) o + f] = r;
    }
  } else if (s - nt_old < 4 * ngen3) {
    u32 const q = s - nt_old;
    u32* d = E.dual + 4 * (size_t) E.nsame3 + q;
    if (*d != DUAL_PENDING) return;
    u32 const key = E.deferred[q];
    u32 const tk = key / 4;
    u32 const n = E.old_dual[key];
    *d = across(n, tk);
  }
}

__global__ __launch_bounds__(BLOCK) void copy_tags(u32 n, TagList L) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  for (u32 t = 0; t < L.n; ++t) {
    u32 w = L.ncomps[t];
    if (L.f64[t]) {
      double const* f = static_cast<double const*>(L.src[t]);
      double* dd = static_cast<double*>(L.dst[t]);
      for (u32 c = 0; c < w; ++c) dd[(size_t) i * w + c] = f[(size_t) i * w + c];
    } else {
      u32 const* f = static_cast<u32 const*>(L.src[t]);
      u32* dd = static_cast<u32*>(L.dst[t]);
      for (u32 c = 0; c < w; ++c) dd[(size_t) i * w + c] = f[(size_t) i * w + c];
    }
  }
}

struct SwapKeepArgs {
  u32 n, ne, vpe, nsame;
  u32 const* soff;
  u32 const* goff;
  u32 const* vent;
  u32* out;
  SwapEdgesOf E;  // tets: the survivors' edges (E.out null: not built)
  u32 const* cd;
  u32 const* ci;
  u32 const* ecd;
  u32 const* eci;
  u32* out_cd;
  u32* out_ci;
};

// concat_verts_of_ents for the survivors of a swap, and the class tags of
// survivors (their own) and of products (their edge's, setup_swap,
// inherit.cpp:158-166): threads [0, n) are entities, [n, n + ne) edges
__global__ __launch_bounds__(BLOCK) void swap_keep(SwapKeepArgs a) {
  u32 t = xcd_block() * BLOCK + threadIdx.x;
  if (t < a.n) {
    u32 o = a.soff[t];
    if (o == a.soff[t + 1]) return;
    for (u32 j = 0; j < a.vpe; ++j) a.out[(size_t) o * a.vpe + j] = a.vent[(size_t) t * a.vpe + j];
    if (a.E.out)
      for (u32 j = 0; j < 6; ++j) a.E.out[(size_t) o * 6 + j] = a.E.soff1[a.E.old[(size_t) t * 6 + j]];
    if (a.out_cd) a.out_cd[o] = a.cd[t];
    if (a.out_ci) a.out_ci[o] = a.ci[t];
    return;
  }
  u32 e = t - a.n;
  if (e >= a.ne) return;
  for (u32 k = a.goff[e]; k < a.goff[e + 1]; ++k) {
    if (a.out_cd) a.out_cd[a.nsame + k] = a.ecd[e];
    if (a.out_ci) a.out_ci[a.nsame + k] = a.eci[e];
  }
}

// a vertex pair as one key (smaller vertex in the high word) and its home
// slot in a table of mask + 1 slots
__host__ __device__ inline unsigned long long swap_pair_key(u32 a, u32 b) {
  return a < b ? ((unsigned long long) a << 32) | b : ((unsigned long long) b << 32) | a;
}
__host__ __device__ inline u32 swap_pair_slot(unsigned long long key, u32 mask) {
  key ^= key >> 29;
  key *= 0xBF58476D1CE4E5B9ull;
  key ^= key >> 32;
  return (u32) key & mask;
}

// swap_common + swap_interior + swap_ents (swap.cpp:27-136)
static u32 swap_common(Mesh& m, Buf<u32>& cand, double const* eq_given = nullptr, SwapCache* cache = nullptr) {
  check_index_space(m);
  u32 ne = count(m, 1);
  unmark_boundary(m, 1, cand.get());
  Buf<double> q(ne);
  Buf<u32> rsz = filled(ne, 0);
  Buf<u32> codes(ne);
  // candidates with a cached choice (their edge's tets unchanged since an
  // earlier swap pass of this shape loop evaluated it) take it: a rejected
  // one leaves the candidates, an accepted one keeps its value, ring size
  // and triangulation; they are counted on the device
  if (cache && cache->ne != ne) cache->clear();
  SwapCache* const cc = cache && cache->ne ? cache : nullptr;
  u32 const* inf = cc ? cc->info.get() : nullptr;
  if (cc) {
    fill(scratch_words() + W_CACHED, 1, 0);
    u32* cp = cand.get();
    double const* cq = cc->q.get();
    double* qo = q.get();
    u32* rs = rsz.get();
    u32* cd = codes.get();
    u32* nacc = scratch_words() + W_CACHED;
    parallel_for(ne, OHX_LAMBDA(u32 i) {
      bool hit = false;
      u32 const f = inf[i];
      if (cp[i] && (f >> 31)) {
        if ((f >> 30) & 1u) {
          qo[i] = cq[i];
          rs[i] = (f >> 16) & 0xffu;
          cd[i] = f & 0xffffu;
          hit = true;
        } else {
          cp[i] = 0;
        }
      }
      unsigned long long const b = __ballot(hit), act = __ballot(1);
      if (b && (u32) __lane_id() == (u32) (__ffsll((long long) act) - 1)) atomicAdd(nacc, (u32) __popcll(b));
    });
  }
  // the candidates to evaluate, compacted (with the cached ones accepted,
  // their count is the reference's early return, swap.cpp:103-106, in one
  // read)
  Buf<u32> coff(ne + 1);
  u32 ncand, ncached = 0;
  {
    u32 const* cp = cand.get();
    auto f = OHX_LAMBDA(u32 i) { return cp[i] && !(inf && (inf[i] >> 31)) ? 1u : 0u; };
    if (cc) {
      exscan_fn(ne, f, coff.get());
      u32 const* ptrs[2] = {coff.get() + ne, scratch_words() + W_CACHED};
      u32 got[2];
      read_many(ptrs, 2, got);
      ncand = got[0];
      ncached = got[1];
    } else {
      ncand = exscan_fn_total(ne, f, coff.get());
    }
  }
  if (!ncand && !ncached) return 0;
  Buf<u32> clist(ncand);
  if (ncand) {
    u32 const* co = coff.get();
    u32* cl = clist.get();
    parallel_for(ne, OHX_LAMBDA(u32 i) {
      if (co[i] != co[i + 1]) cl[co[i]] = i;
    });
  }
  coff.release();
  make_swap_tables();
  Up const& u = ask_up(m, 1, 3);
  u32 const* ev = ask_down(m, 1, 0);
  u32 const* vt = ask_down(m, 3, 0);
  Buf<double> eq;
  if (!eq_given) eq = mesh_qualities(m);
  double const* eqp = eq_given ? eq_given : eq.get();
  u32 tag = next_any_tag();
  if (ncand) {
    // compulsory bytes: the edges' candidate flags, vertices and outputs,
    // the tets around them (uses + directions), their qualities and
    // vertices, and the vertex coordinates (and metrics)
    double const* mts = metric_of(m);
    double prof_bytes = (double) ne * 24 + (double) u.adj.size() * 8 + (double) count(m, 3) * 24 +
                        (double) count(m, 0) * (mts ? 72 : 24);
    ScreenStats sst = screen_stats();
    dim3 grid((unsigned) ((ncand + SW_EDGES - 1) / SW_EDGES));
    auto go = [&](auto kernel, char const* kn, hipEvent_t ea, hipEvent_t eb) {
      if (ea) {
        hipExtLaunchKernelGGL(kernel, grid, dim3(BLOCK), 0, stream(), ea, eb, 0, ncand, clist.get(), u.offsets.get(),
                              u.adj.get(), u.dirs.get(), ev, vt, coords(m), mts, eqp, cand.get(), q.get(),
                              rsz.get(), codes.get(), scratch_words(), tag, dev_error_word(), sst);
        check_launch(kn);
      } else {
        launch_timed(kn, prof_bytes, kernel, grid, dim3(BLOCK), ncand, clist.get(), u.offsets.get(), u.adj.get(),
                     u.dirs.get(), ev, vt, coords(m), mts, eqp, cand.get(), q.get(), rsz.get(), codes.get(),
                     scratch_words(), tag, dev_error_word(), sst);
      }
    };
    // in the mixed mode the FP32 screen is a variant chosen online: it pays
    // only where it rejects enough candidates (most swap candidates improve
    // their cavity and need their exact FP64 value anyway)
    static thread_local Tuner tuner("swap_quality_3d_mixed", {0, 1});
    auto run = [&](u32 screen, hipEvent_t ea, hipEvent_t eb) {
      if (screen)
        mts ? go(swap_quality_k<true, true>, "swap_quality_mixed", ea, eb)
            : go(swap_quality_k<false, true>, "swap_quality_mixed", ea, eb);
      else
        mts ? go(swap_quality_k<true, false>, "swap_quality_f64", ea, eb)
            : go(swap_quality_k<false, false>, "swap_quality_f64", ea, eb);
    };
    if (g_mixed)
      tuned(tuner, ncand, run);
    else
      run(0, nullptr, nullptr);
  }
  eq.release();
  // the evaluated candidates' choices go into the cache (this mesh's
  // numbering; the swap below carries the cache to the new mesh)
  if (cache) {
    if (!cc) {
      cache->ne = ne;
      cache->q = Buf<double>(ne);
      cache->info = filled(ne, 0);
    }
    if (ncand) {
      u32 const* cl = clist.get();
      u32 const* cp = cand.get();
      double const* qi = q.get();
      u32 const* rs = rsz.get();
      u32 const* cd = codes.get();
      double* cq = cache->q.get();
      u32* ci = cache->info.get();
      parallel_for(ncand, OHX_LAMBDA(u32 k) {
        u32 const i = cl[k];
        ci[i] = (1u << 31) | (cp[i] ? (1u << 30) | (rs[i] << 16) | cd[i] : 0u);
        cq[i] = qi[i];
      });
    }
  }
  // the early return after the qualities (swap.cpp:119-124) is the totals
  // read below: the quality kernel clears every rejected candidate, so
  // with none accepted the set is empty and no element is generated
  // a first batch of indset rounds with no host check, every dimension's
  // product and survivor offsets, then one read of the convergence flag
  // and all the totals (a set not settled in the first batch goes on from
  // there with the checked loop, same rounds, same result)
  u32 const* flag = nullptr;
  Buf<u32> indset = indset_rounds(m, 1, cand.get(), q.get(), first_batch_rounds(), &flag, true);
  u32 dlo = m.full ? 1 : m.dim;
  Buf<u32> goffs[4], soffs[4];
  u32 tots[8];
  // the faces that survive are those holding no swapped edge; a face holds
  // an edge exactly when it holds both of the edge's vertices, so when the
  // face-to-edge map is not derived (it is needed for nothing else here) the
  // swapped edges' vertex pairs go into a small hash table and each face
  // looks its three vertex pairs up, instead of hashing every edge of the
  // mesh to build down(2,1)
  bool const faces_by_pairs = m.dim == 3 && m.full && !m.down[2][1];
  u32 pcap = 64;
  while (pcap < 4 * ((size_t) ncand + ncached)) pcap *= 2;
  Buf<unsigned long long> pairs(faces_by_pairs ? pcap : 0);
  auto offsets_and_totals = [&] {
    u32 const* is = indset.get();
    u32 const* rs = rsz.get();
    u32 const* pt[8];
    u32 k = 0;
    pt[k++] = flag;
    if (faces_by_pairs) {
      unsigned long long* tb = pairs.get();
      u32 const mask = pcap - 1;
      parallel_for(pcap, OHX_LAMBDA(u32 i) { tb[i] = ~0ull; });
      u32* err = dev_error_word();
      parallel_for(ne, OHX_LAMBDA(u32 e) {
        if (!is[e]) return;
        unsigned long long key = swap_pair_key(ev[2 * (size_t) e], ev[2 * (size_t) e + 1]);
        u32 h = swap_pair_slot(key, mask);
        for (u32 probe = 0; probe <= mask; ++probe, h = (h + 1) & mask)
          if (atomicCAS(&tb[h], ~0ull, key) == ~0ull) return;
        *err = ERR_REFLECT_NOT_FOUND;
      });
    }
    for (u32 d = dlo; d <= m.dim; ++d) {
      u32 n = count(m, d);
      // products per swapped edge (edge_topo_offset, swap_topology.cpp:16-44)
      // and survivors: entities adjacent to no swapped edge (swap.cpp:41-45)
      goffs[d] = Buf<u32>(ne + 1);
      soffs[d] = Buf<u32>(n + 1);
      exscan_fn(ne, OHX_LAMBDA(u32 i) {
        return is[i] ? count_swap_ents(rs[i], d, k_sw_nint[rs[i]], k_sw_sizes[rs[i]]) : 0u;
      }, goffs[d].get());
      if (d == 2 && faces_by_pairs) {
        u32 const* fv = ask_down(m, 2, 0);
        unsigned long long const* tb = pairs.get();
        u32 const mask = pcap - 1;
        exscan_fn_heavy(n, OHX_LAMBDA(u32 i) {
          u32 const* v = fv + 3 * (size_t) i;
          for (u32 j = 0; j < 3; ++j) {
            unsigned long long key = swap_pair_key(v[j], v[(j + 1) % 3]);
            for (u32 probe = 0, h = swap_pair_slot(key, mask); probe <= mask; ++probe, h = (h + 1) & mask) {
              unsigned long long t = tb[h];
              if (t == key) return 0u;
              if (t == ~0ull) break;
            }
          }
          return 1u;
        }, soffs[d].get());
      } else {
        u32 lpe = h_down_deg[d][1];
        u32 const* le = ask_down(m, d, 1);
        exscan_fn_heavy(n, OHX_LAMBDA(u32 i) {
          for (u32 j = 0; j < lpe; ++j)
            if (is[le[(size_t) i * lpe + j]]) return 0u;
          return 1u;
        }, soffs[d].get());
      }
      pt[k++] = goffs[d].get() + ne;
      pt[k++] = soffs[d].get() + n;
    }
    read_many(pt, k, tots);
  };
  offsets_and_totals();
  if (tots[0]) {  // not settled in the first batch: the checked loop from there
    indset = find_indset_resume(m, 1, q.get(), std::move(indset), true);
    flag = nullptr;
    offsets_and_totals();
  }
  if (g_batch) batch_note_changed(m, 1, indset.get());
  cand.release();
  q.release();
  // every swapped edge generates elements, so none generated: nothing swapped
  if (!tots[1 + 2 * (m.dim - dlo)]) return 0;
  u32 total = g_verbose ? sum_u32(indset.get(), ne) : 0u;
  u32 nv = count(m, 0);
  Mesh mo(m.dim, m.full);
  set_ents(mo, 0, nv, Buf<u32>());
  {
    // copy_tags (tag.cpp): the vertices are unchanged, one launch for all tags
    TagList L;
    for (auto& tp : m.tags[0]) {
      Tag& t = *tp;
      void* dst;
      if (t.type == TagType::F64)
        dst = add_tag_f64(mo, 0, t.name.c_str(), t.ncomps, Buf<double>((size_t) nv * t.ncomps), t.transfer).f.get();
      else
        dst = add_tag_u32(mo, 0, t.name.c_str(), t.ncomps, Buf<u32>((size_t) nv * t.ncomps), t.transfer).u.get();
      tag_list_add(L, t, dst);
    }
    if (L.n) {
      hipLaunchKernelGGL(copy_tags, dim3(nblocks(nv)), dim3(BLOCK), 0, stream(), nv, L);
      check_launch("copy_tags");
    }
  }
  u32 const* is = indset.get();
  bool has_class = m.full && find_tag(m, 0, "class_dim") != nullptr;
  bool has_id = has_class && find_tag(m, 0, "class_id") != nullptr;
  // the tets' edges of the new mesh, from the old map (the old one exists:
  // the rings above read up(1,3)); the next pass asks for them first
  bool const edges_of_tets = m.dim == 3 && dlo == 1;
  Buf<u32> d31, dual_new, face_prod, deferred;
  // every dimension's products in one launch (the ring of each swapped edge
  // found once), then the survivors dimension by dimension
  Buf<u32> vouts[4];
  u32* pouts[4] = {nullptr, nullptr, nullptr, nullptr};
  double prof_bytes = (double) ne * 12;
  for (u32 d = dlo; d <= m.dim; ++d) {
    u32 ngen = tots[1 + 2 * (d - dlo)], nsame = tots[2 + 2 * (d - dlo)];
    vouts[d] = Buf<u32>((size_t) (nsame + ngen) * (d + 1));
    pouts[d] = vouts[d].get() + (size_t) nsame * (d + 1);
    prof_bytes += (double) ngen * (d + 1) * 4;
  }
  SwapEdgesOf E{};
  u32 const nt_old = count(m, 3), ngen3 = tots[1 + 2 * (3 - dlo)], nsame3 = tots[2 + 2 * (3 - dlo)];
  if (edges_of_tets) {
    d31 = Buf<u32>((size_t) (nsame3 + ngen3) * 6);
    E.out = d31.get();
    E.old = ask_down(m, 3, 1);
    E.soff1 = soffs[1].get();
    E.goff1 = goffs[1].get();
    E.nsame1 = tots[2];
    E.nsame3 = nsame3;
    E.err = dev_error_word();
    if (m.dual) {
      // the dual this pass's sliver marks read, carried to the next mesh
      dual_new = Buf<u32>((size_t) (nsame3 + ngen3) * 4);
      face_prod = Buf<u32>((size_t) nt_old * 4);
      deferred = Buf<u32>((size_t) ngen3 * 4);
      E.dual = dual_new.get();
      E.old_dual = m.dual.get();
      E.soff3 = soffs[3].get();
      E.face_prod = face_prod.get();
      E.deferred = deferred.get();
    }
  }
  launch_timed("swap_cavity", prof_bytes, swap_products_k, dim3(nblocks(ne)), dim3(BLOCK), ne, is,
               goffs[1].get(), goffs[2].get(), goffs[3].get(), u.offsets.get(), u.adj.get(), u.dirs.get(), ev, vt,
               codes.get(), pouts[1], pouts[2], pouts[3], E);
  if (E.dual) {
    hipLaunchKernelGGL(swap_dual_k, dim3(nblocks((size_t) nt_old + (size_t) ngen3 * 4)), dim3(BLOCK), 0, stream(),
                       nt_old, ngen3, E);
    check_launch("swap_dual");
  }
  for (u32 d = dlo; d <= m.dim; ++d) {
    u32 n = count(m, d);
    u32 vpe = d + 1;
    Buf<u32>& goff = goffs[d];
    Buf<u32>& soff = soffs[d];
    u32 ngen = tots[1 + 2 * (d - dlo)], nsame = tots[2 + 2 * (d - dlo)];
    Buf<u32> vout = std::move(vouts[d]);
    Buf<u32> cdo, cio;
    SwapKeepArgs a{};
    if (d == 3) a.E = E;
    a.n = n;
    a.ne = ne;
    a.vpe = vpe;
    a.nsame = nsame;
    a.soff = soff.get();
    a.goff = goff.get();
    a.vent = ask_down(m, d, 0);
    a.out = vout.get();
    if (has_class) {
      cdo = Buf<u32>(nsame + ngen);
      a.cd = u32_tag(m, d, "class_dim");
      a.ecd = u32_tag(m, 1, "class_dim");
      a.out_cd = cdo.get();
      if (has_id) {
        cio = Buf<u32>(nsame + ngen);
        a.ci = u32_tag(m, d, "class_id");
        a.eci = u32_tag(m, 1, "class_id");
        a.out_ci = cio.get();
      }
    }
    hipLaunchKernelGGL(swap_keep, dim3(nblocks(n + ne)), dim3(BLOCK), 0, stream(), a);
    check_launch("swap_keep");
    set_ents(mo, d, nsame + ngen, std::move(vout));
    if (has_class) {
      add_tag_u32(mo, d, "class_dim", 1, std::move(cdo));
      if (has_id) add_tag_u32(mo, d, "class_id", 1, std::move(cio));
    }
    if (d == m.dim) swap_transfer(m, mo, goff.get(), soff.get(), nsame, ngen);
    if (d == m.dim && cache && eq_given) {
      // the new mesh's element qualities: a survivor's is its old one, a
      // product's is evaluated (the same kernel mesh_qualities runs)
      Buf<double> neq = element_qualities(d, ngen, mo.down[d][0].get() + (size_t) nsame * vpe, coords(m),
                                          count(m, 0), metric_of(m));
      Buf<double> all((size_t) nsame + ngen);
      double* ap = all.get();
      double const* np = neq.get();
      u32 const* so = soff.get();
      double const* oq = eq_given;
      parallel_for(n, OHX_LAMBDA(u32 t) {
        u32 const o = so[t];
        if (o != so[t + 1]) ap[o] = oq[t];
      });
      if (ngen) d2d(ap + nsame, np, ngen);
      cache->eq = std::move(all);
    }
  }
  if (d31) {
    // the edge-to-tet adjacency from the old one (the rings above read it)
    UpPatchCounts pc;
    mo.up[1][3] = up_patch(6, u, ne, soffs[1].get(), tots[2], tots[1], soffs[3].get(), tots[6], tots[5], d31.get(),
                           cache ? &pc : nullptr);
    mo.down[3][1] = std::move(d31);
    if (dual_new) mo.dual = std::move(dual_new);
    if (cache) {
      // the cache in the new numbering: a surviving edge that kept all its
      // tets and gained none keeps its entry
      u32 const ne_new = tots[2] + tots[1];
      Buf<double> nq(ne_new);
      Buf<u32> ninfo = filled(ne_new, 0);
      u32 const* so = soffs[1].get();
      u32 const* full = pc.full.get();
      u32 const* cnt = pc.cnt.get();
      u32 const* cnt_s = pc.cnt_s.get();
      u32 const* ci = cache->info.get();
      double const* cq = cache->q.get();
      u32* ni = ninfo.get();
      double* nqp = nq.get();
      parallel_for(ne, OHX_LAMBDA(u32 e) {
        u32 const n = so[e];
        if (n == so[e + 1] || !full[n] || cnt[n] != cnt_s[n]) return;
        ni[n] = ci[e];
        nqp[n] = cq[e];
      });
      cache->ne = ne_new;
      cache->q = std::move(nq);
      cache->info = std::move(ninfo);
    }
  } else if (cache) {
    cache->clear();
  }
  say("swapped", total, 1);
  overwrite(m, std::move(mo));
  return 1;
}

u32 swap_slivers(Mesh& m, double good_qual, u32 nlayers, double const* eq, double const* gate, SwapCache* cache) {
  OHX_REQUIRE(m.dim == 3, "swap_slivers needs a 3D mesh");
  // the pass reads up(1,3) (the edge marks below, the rings, the quality
  // kernel); derived first, it also gives the sliver layers' element dual
  // from the rings (dual_edges_k) instead of a hash table of the faces
  (void) ask_up(m, 1, 3);
  Buf<u32> sl = mark_slivers(m, good_qual, nlayers, eq, gate);
  Buf<u32> c = mark_down(m, m.dim, 1, sl.get());
  sl.release();
  if (g_batch) batch_mask(m, 1, c.get());
  return swap_common(m, c, eq, cache);
}

// ---------------------------------------------------------------- adapt
// mesh_adapt (adapt.cpp:19-94)
namespace {
struct Ctx {
  u32 ops = 0, max_ops = 0;
  AdaptStats st;
  SwapCache swap_cache;  // the shape loop's swap choices (cleared by other passes)
  SizeCodeCache size_codes;  // the collapse loop's size codes (cleared by other passes)
  // passes that changed the mesh: events around them on the stream, read
  // once at the end (a pass's time is the stream's, from the end of the work
  // before it to the end of its own, host waits included)
  struct Timed {
    hipEvent_t a, b;
    double tin;
  };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> pool;
  hipEvent_t event() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    OHX_CHECK(hipEventCreate(&e));
    return e;
  }
  void finish() {
    for (Timed& t : timed) {
      float ms = 0;
      OHX_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
      st.npass += 1;
      st.nelems += t.tin;
      st.tpass += ms * 1e-3;
      pool.push_back(t.a);
      pool.push_back(t.b);
    }
    timed.clear();
    for (hipEvent_t e : pool) (void) hipEventDestroy(e);
    pool.clear();
  }
};
// adapt_summary (adapt.cpp:22-34) only prints: it reads the mesh and changes
// nothing, so it runs only when the per-pass lines are printed
void adapt_summary(Mesh& m) {
  if (!g_verbose) return;
  u32 nel = count(m, m.dim);
  double mq = mesh_min_quality(m);
  u32 ne = count(m, 1);
  Buf<double> s = measure_edges(m);
  double lo = min_f64(s.get(), ne, 1), hi = max_f64(s.get(), ne, 1);
  std::printf("%10u elements, min quality %.0f%%, metric range %.2f - %.2f\n", nel, mq * 100.0, lo, hi);
}
bool incr_op(Mesh& m, Ctx& c) {  // adapt.cpp:36-45
  if (c.ops > c.max_ops) return false;
  ++c.ops;
  adapt_summary(m);
  return true;
}
enum Kind { REFINE, COARSEN, SWAP, COARSEN_SLIVERS };
u32 run_pass(Mesh& m, Ctx& c, Kind k, double a, double b, u32 n, double const* eq = nullptr,
             double const* gate = nullptr) {
  double tin = count(m, m.dim);
  static bool const trace = std::getenv("OHX_TRACE_DERIVE") && std::getenv("OHX_TRACE_DERIVE")[0] == '1';
  if (trace) std::fprintf(stderr, "pass %s n=%.0f\n", k == REFINE ? "refine" : k == COARSEN ? "coarsen" : k == SWAP ? "swap" : "csliv", tin);
  hipEvent_t e0 = c.event(), e1 = c.event();
  OHX_CHECK(hipEventRecord(e0, stream()));
  u32 ok = 0;
  if (k != SWAP) c.swap_cache.clear();
  if (k != COARSEN) c.size_codes.clear();
  switch (k) {
    case REFINE: ok = refine_by_size(m, a); break;
    case COARSEN: ok = coarsen_by_size(m, a, b, &c.size_codes); break;
    case SWAP: ok = swap_slivers(m, a, n, eq, gate, &c.swap_cache); break;
    case COARSEN_SLIVERS: ok = coarsen_slivers(m, a, n, eq); break;
  }
  OHX_CHECK(hipEventRecord(e1, stream()));
  if (ok)
    c.timed.push_back(Ctx::Timed{e0, e1, tin});
  else {
    c.pool.push_back(e0);
    c.pool.push_back(e1);
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
    // the element qualities of this mesh, evaluated once: its minimum, the
    // sliver marks and the comparisons of the swap and collapse passes all
    // read them (a pass that changes nothing leaves them valid for the next)
    u32 const nel = count(m, m.dim);
    Buf<double> eq = c.swap_cache.eq ? std::move(c.swap_cache.eq) : mesh_qualities(m);
    if (m.dim == 3) {
      // the minimum (over elements 1..n-1, as the serial doubles_min)
      // reaches the host with the swap pass's first wait instead of a wait
      // of its own: the pass's sliver marks are gated on it on the device,
      // so when it is at least good_qual the pass marks nothing, changes
      // nothing, and the loop stops where the reference's stops
      Buf<double> mq(1);
      min_f64_async(eq.get(), nel, 1, mq.get());
      u32 const swapped = run_pass(m, c, SWAP, good_qual, 0, nlayers, eq.get(), mq.get());
      double const prev = min_f64_async_result();
      if (prev >= good_qual) {
        OHX_REQUIRE(!swapped, "shape loop: a swap pass ran on a mesh of good quality");
        break;
      }
      if (swapped) {
        if (!incr_op(m, c)) aborted = true;
        continue;
      }
    } else {
      double prev = nel > 1 ? min_f64(eq.get(), nel, 1) : 1.7976931348623157e308;
      if (prev >= good_qual) break;
    }
    if (run_pass(m, c, COARSEN_SLIVERS, good_qual, 0, nlayers, eq.get())) {
      if (!incr_op(m, c)) aborted = true;
      continue;
    }
    if (g_verbose) std::printf("warning: shape quality could not be satisfied\n");
    break;
  }
  result = aborted ? -1 : (c.ops > 0 ? 1 : 0);
  sync();
  c.finish();
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
    double mq = min_element_quality(dim, ne, ev, xp, nv, metric_of(m));
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
