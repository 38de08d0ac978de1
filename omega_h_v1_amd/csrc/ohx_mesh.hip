// Mesh container, lazy topology derivation, quality kernels and marks.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ohx_geom.hpp"
#include "ohx_mesh.hpp"
#include "ohx_prims.hpp"
#include "ohx_tables.hpp"

namespace ohx {

// ------------------------------------------------------------- bookkeeping
void set_ents(Mesh& m, u32 d, u32 n, Buf<u32>&& verts) {
  OHX_REQUIRE(!m.down[d][0] || d == 0, "entities of this dimension were already set");
  m.counts[d] = n;
  if (d) m.down[d][0] = std::move(verts);
}

bool has_dim(Mesh& m, u32 d) { return d == 0 || bool(m.down[d][0]); }

u32 count(Mesh& m, u32 d) {
  if (d > m.dim) return 0;
  if (d) ask_down(m, d, 0);
  return m.counts[d];
}

void overwrite(Mesh& m, Mesh&& w) {
  m.dim = w.dim;
  m.full = w.full;
  for (u32 i = 0; i < 4; ++i) {
    m.counts[i] = w.counts[i];
    m.tags[i] = std::move(w.tags[i]);
    for (u32 j = 0; j < 4; ++j) {
      m.down[i][j] = std::move(w.down[i][j]);
      m.up[i][j] = std::move(w.up[i][j]);
      m.star[i][j] = std::move(w.star[i][j]);
    }
  }
  m.dual = std::move(w.dual);
  for (u32 i = 0; i < 4; ++i) {
    m.low_table[i] = std::move(w.low_table[i]);
    m.low_mask[i] = w.low_mask[i];
  }
}

Tag* find_tag(Mesh& m, u32 d, char const* name) {
  for (auto& t : m.tags[d])
    if (t->name == name) return t.get();
  return nullptr;
}

static Tag& push_tag(Mesh& m, u32 d, char const* name, TagType type, u32 ncomps, int transfer) {
  OHX_REQUIRE(!find_tag(m, d, name), "tag already exists");
  auto t = std::make_unique<Tag>();
  t->name = name;
  t->type = type;
  t->ncomps = ncomps;
  t->transfer = transfer;
  m.tags[d].push_back(std::move(t));
  return *m.tags[d].back();
}

Tag& add_tag_u32(Mesh& m, u32 d, char const* name, u32 ncomps, Buf<u32>&& data, int transfer) {
  Tag& t = push_tag(m, d, name, TagType::U32, ncomps, transfer);
  t.u = std::move(data);
  return t;
}

Tag& add_tag_f64(Mesh& m, u32 d, char const* name, u32 ncomps, Buf<double>&& data, int transfer) {
  Tag& t = push_tag(m, d, name, TagType::F64, ncomps, transfer);
  t.f = std::move(data);
  return t;
}

void remove_tag(Mesh& m, u32 d, char const* name) {
  auto& v = m.tags[d];
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i]->name == name) {
      v.erase(v.begin() + i);
      return;
    }
  fail("removing a tag that does not exist", __FILE__, __LINE__);
}

double const* coords(Mesh& m) {
  Tag* t = find_tag(m, 0, "coordinates");
  OHX_REQUIRE(t && t->type == TagType::F64, "mesh has no coordinates");
  return t->f.get();
}

double const* metric_of(Mesh& m) {
  Tag* t = find_tag(m, 0, "metric");
  if (!t) return nullptr;
  OHX_REQUIRE(t->type == TagType::F64 && t->ncomps == 6, "metric field must have 6 FP64 components");
  return t->f.get();
}

double const* size_of(Mesh& m) {
  if (metric_of(m)) return nullptr;
  Tag* st = find_tag(m, 0, "adapt_size");
  OHX_REQUIRE(st && st->type == TagType::F64, "mesh has neither an adapt_size nor a metric field");
  return st->f.get();
}

u32 const* u32_tag(Mesh& m, u32 d, char const* name) {
  Tag* t = find_tag(m, d, name);
  OHX_REQUIRE(t && t->type == TagType::U32, "missing integer tag");
  return t->u.get();
}

// ---------------------------------------------------------------- quality
// One thread per element: gather the vertex coordinates, evaluate the
// reference's quality formula in FP64 (quality.cpp:13-45).

template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void elem_quality_f64(u32 n, u32 const* __restrict__ ev,
                                                          double const* __restrict__ x,
                                                          double const* __restrict__ mt,
                                                          double* __restrict__ out) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  out[i] = elem_q<DIM>(i, ev, x, mt);
}

// Fused quality + minimum over elements begin..n-1 (no per-element store).
template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void elem_min_quality_f64(u32 n, u32 begin,
                                                              u32 const* __restrict__ ev,
                                                              double const* __restrict__ x,
                                                              double const* __restrict__ mt,
                                                              double* __restrict__ part) {
  __shared__ double lds[BLOCK / 64];
  double v = 1.7976931348623157e308;
  for (u32 i = begin + xcd_block() * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    double q = elem_q<DIM>(i, ev, x, mt);
    if (q < v) v = q;
  }
#pragma unroll
  for (u32 d = 32; d >= 1; d >>= 1) {
    double o = __shfl_down(v, d, 64);
    if (o < v) v = o;
  }
  if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = lds[0];
    for (u32 w = 1; w < BLOCK / 64; ++w)
      if (lds[w] < r) r = lds[w];
    part[blockIdx.x] = r;
  }
}

Buf<double> element_qualities(u32 dim, u32 n, u32 const* ev, double const* x, u32 nverts, double const* mt) {
  Buf<double> q(n);
  if (!n) return q;
  // compulsory bytes: connectivity, each vertex's coordinates once, output
  double prof_bytes = (double) n * (dim + 1) * 4 + (double) nverts * 24 + (double) n * 8;
  if (dim == 3)
    launch_timed("elem_quality_f64", prof_bytes, elem_quality_f64<3>, dim3(nblocks(n)), dim3(BLOCK), n, ev, x, mt, q.get());
  else
    launch_timed("elem_quality_f64", prof_bytes, elem_quality_f64<2>, dim3(nblocks(n)), dim3(BLOCK), n, ev, x, mt, q.get());
  return q;
}

Buf<double> mesh_qualities(Mesh& m) {
  return element_qualities(m.dim, count(m, m.dim), ask_down(m, m.dim, 0), coords(m), count(m, 0), metric_of(m));
}

// min over elements 1..n-1: the serial build's doubles_min starts at 1
// (doubles.cpp:111-118), and mesh_min_quality goes through it
double min_element_quality(u32 dim, u32 n, u32 const* ev, double const* x, u32 nverts, double const* mt) {
  if (n <= 1) return 1.7976931348623157e308;
  u32 nb = std::min<u32>(nblocks(n - 1), 1024);
  Buf<double> part(nb);
  {
    double prof_bytes = (double) n * (dim + 1) * 4 + (double) nverts * 24 + (double) nb * 8;
    if (dim == 3)
      launch_timed("elem_min_quality_f64", prof_bytes, elem_min_quality_f64<3>, dim3(nb), dim3(BLOCK), n, 1u, ev, x, mt, part.get());
    else
      launch_timed("elem_min_quality_f64", prof_bytes, elem_min_quality_f64<2>, dim3(nb), dim3(BLOCK), n, 1u, ev, x, mt, part.get());
  }
  return min_f64(part.get(), nb, 0);
}

double mesh_min_quality(Mesh& m) {
  return min_element_quality(m.dim, count(m, m.dim), ask_down(m, m.dim, 0), coords(m), count(m, 0), metric_of(m));
}

// measure_edge (size.cpp:57-88): |x1 - x0| / ((h0 + h1) / 2)
__global__ __launch_bounds__(BLOCK) void measure_edge_f64(u32 n, u32 const* __restrict__ ev,
                                                          double const* __restrict__ x,
                                                          double const* __restrict__ h,
                                                          double const* __restrict__ mt,
                                                          double* __restrict__ out) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  out[i] = edge_size(ev[2 * i], ev[2 * i + 1], x, h, mt);
}

Buf<double> measure_edges(Mesh& m) {
  u32 ne = count(m, 1);
  double const* h = size_of(m);
  double const* mt = metric_of(m);
  Buf<double> out(ne);
  if (ne) {
    double prof_bytes = (double) ne * 16 + (double) count(m, 0) * (mt ? 72 : 32);
    launch_timed("measure_edge_f64", prof_bytes, measure_edge_f64, dim3(nblocks(ne)), dim3(BLOCK), ne,
                       ask_down(m, 1, 0), coords(m), h, mt, out.get());
  }
  return out;
}

// --------------------------------------------------------- up_from_down
// Upward adjacency as a stable counting sort of uses by low entity
// (invert_map.cpp:72-136 + up_from_down.cpp:9-47): per low, the uses in
// increasing use index (= high * deg + direction).  The counting atomics
// also hand each use a slot in its bucket; each bucket is then sorted, which
// makes the result exact whatever order the atomics ran in.
// COMBINE: a wave combines its adds per distinct low (one atomic each),
// which pays where the uses of a low sit close together in the down map (a
// freshly generated mesh) and costs a loop over the wave's distinct lows
// where they do not (a mesh after many adapt passes)
template <bool COMBINE>
__global__ __launch_bounds__(BLOCK) void upfd_count(u32 nuses, u32 const* __restrict__ in,
                                                    u32* __restrict__ counts, u32* __restrict__ slot) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (COMBINE) {
    bool live = i < nuses;
    u32 s = wave_slot_add(counts, live ? in[i] : 0u, live);
    if (live) slot[i] = s;
  } else if (i < nuses) {
    slot[i] = atomicAdd(&counts[in[i]], 1u);
  }
}

// every use to its slot (no atomics); the use that took slot 0 of a bucket
// clears that bucket's counter, so all counters are zero again afterwards
__global__ __launch_bounds__(BLOCK) void upfd_fill(u32 nuses, u32 const* __restrict__ in,
                                                   u32 const* __restrict__ slot, u32 const* __restrict__ offs,
                                                   u32* __restrict__ counts, u32* __restrict__ out) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= nuses) return;
  u32 d = in[i];
  u32 j = slot[i];
  out[offs[d] + j] = i;
  if (j == 0) counts[d] = 0;
}

// G lanes per low entity: each use's final slot is its rank among the
// bucket's uses (use indices are distinct, so the rank is the count of
// smaller ones); the bucket is read from the fill buffer and written sorted,
// split into (high, direction).  Lanes of a group read the same word at a
// time (one transaction per group), and a bucket costs k^2 / G compares.
template <u32 G>
__global__ __launch_bounds__(BLOCK) void upfd_rank_split(u32 nlows, u32 deg, u32 const* __restrict__ offs,
                                                         u32 const* __restrict__ tmp, u32* __restrict__ adj,
                                                         u32* __restrict__ dirs) {
  size_t gid = (size_t) xcd_block() * BLOCK + threadIdx.x;
  size_t grp = gid / G;
  u32 lane = (u32) (gid % G);
  if (grp >= nlows) return;
  u32 f = offs[grp], e = offs[grp + 1];
  for (u32 j = f + lane; j < e; j += G) {
    u32 v = tmp[j];
    u32 r = 0;
    for (u32 k = f; k < e; ++k) r += tmp[k] < v ? 1u : 0u;
    u32 h = v / deg;
    adj[f + r] = h;
    dirs[f + r] = v - h * deg;
  }
}

// Ascending sort of one value per lane across a segment of S lanes
// (bitonic network over lane permutes within the segment).
template <u32 S>
__device__ inline u32 segment_sort(u32 v, u32 lane) {
#pragma unroll
  for (u32 k = 2; k <= S; k <<= 1)
#pragma unroll
    for (u32 j = k >> 1; j > 0; j >>= 1) {
      u32 o = (u32) __shfl_xor((int) v, (int) j, (int) S);
      bool const keep_min = ((lane & j) == 0) == ((lane & k) == 0);
      v = keep_min ? (o < v ? o : v) : (o > v ? o : v);
    }
  return v;
}

// S lanes per low entity, the bucket held in the lanes: lane j of a segment
// holds the bucket's j-th use, and each use's rank is counted against the
// whole bucket read across the segment with lane permutes (no memory reads
// in the loop).  Every lane of the wave takes part in those permutes.  A
// bucket longer than S is sorted in chunks of S uses in place in the fill
// buffer (each chunk by the segment's lanes), and a use's rank is then its
// place in its own chunk plus, for every other chunk, the count of smaller
// uses found by binary search: k (k/S) log S reads instead of the k^2 of
// counting against the whole bucket.  (The fill buffer only holds each
// bucket's uses in some order, so sorting it in place changes no result.)
template <u32 S>
__global__ __launch_bounds__(BLOCK) void upfd_rank_seg(u32 nlows, u32 deg, u32 const* __restrict__ offs,
                                                       u32* __restrict__ tmp, u32* __restrict__ adj,
                                                       u32* __restrict__ dirs) {
  size_t gid = (size_t) xcd_block() * BLOCK + threadIdx.x;
  size_t grp = gid / S;
  u32 const lane = (u32) (gid % S);
  u32 const seg0 = (threadIdx.x & 63u) & ~(S - 1);
  bool const live = grp < nlows;
  u32 f = 0, e = 0;
  if (live) {
    f = offs[grp];
    e = offs[grp + 1];
  }
  u32 const k = e - f;
  // use indices are below 2^32 - 1, so the padding never counts as smaller
  u32 const v = (live && lane < k) ? tmp[f + lane] : INVALID;
  u32 r = 0;
#pragma unroll 8
  for (u32 t = 0; t < S; ++t) {
    u32 o = (u32) __shfl((int) v, (int) (seg0 + t), 64);
    r += o < v ? 1u : 0u;
  }
  if (!live) return;
  if (k <= S) {
    if (lane < k) {
      u32 h = v / deg;
      adj[f + r] = h;
      dirs[f + r] = v - h * deg;
    }
    return;
  }
  // the segment's lanes run these loops together (same bucket, same trip
  // counts), so the permutes of segment_sort stay within active lanes
  for (u32 c = f; c < e; c += S) {
    u32 const idx = c + lane;
    u32 w = idx < e ? tmp[idx] : INVALID;
    w = segment_sort<S>(w, lane);
    if (idx < e) tmp[idx] = w;
  }
  __threadfence_block();
  for (u32 c = f; c < e; c += S) {
    u32 const idx = c + lane;
    if (idx >= e) continue;
    u32 const w = tmp[idx];
    u32 rr = lane;
    for (u32 c2 = f; c2 < e; c2 += S) {
      if (c2 == c) continue;
      u32 lo = c2, hi = c2 + S < e ? c2 + S : e;
      while (lo < hi) {
        u32 mid = (lo + hi) >> 1;
        if (tmp[mid] < w) lo = mid + 1;
        else hi = mid;
      }
      rr += lo - c2;
    }
    u32 h = w / deg;
    adj[f + rr] = h;
    dirs[f + rr] = w - h * deg;
  }
}

// The bucket counters: one buffer per host thread, all zero between uses
// (upfd_fill clears every bucket it fills), so a call needs no clearing
// launch.  dirty marks a call that did not get as far as the fill.
namespace {
struct ZeroCounters {
  u32* p = nullptr;
  size_t cap = 0;
  bool dirty = false;
};
ZeroCounters& zero_counters() {
  static thread_local ZeroCounters z;
  return z;
}
u32* take_zeroed_counters(u32 n) {
  ZeroCounters& z = zero_counters();
  if (n > z.cap) {
    // stream-ordered release: the arena reuses the block only after the
    // stream's earlier work
    if (z.p) dev_free(z.p);
    size_t cap = z.cap * 2 > n ? z.cap * 2 : n;
    if (cap > 0xffffffffull) cap = 0xffffffffull;
    z.cap = cap < ((size_t) 1 << 16) ? ((size_t) 1 << 16) : cap;
    z.p = static_cast<u32*>(dev_alloc(z.cap * sizeof(u32)));
    z.dirty = true;
  }
  if (z.dirty) {
    fill(z.p, (u32) z.cap, 0);
    z.dirty = false;
  }
  return z.p;
}
}  // namespace

void up_from_down(u32 hd, u32 ld, u32 nhighs, u32 nlows, u32 const* lows_of_highs, Up& u) {
  u32 deg = h_down_deg[hd][ld];
  if ((size_t) nhighs * deg >= ((size_t) 1 << 32)) fail("up_from_down: more than 2^32 uses", __FILE__, __LINE__);
  u32 nuses = nhighs * deg;
  u32* counts = take_zeroed_counters(nlows);
  Buf<u32> slot(nuses);
  if (nuses) {
    zero_counters().dirty = true;
    // compulsory bytes: the down map in, each counter read and written
    // compulsory bytes: the down map in, slots out, each counter read and
    // written; plain or wave-combined adds, chosen online per (high, low)
    // dimension pair and size class
    static thread_local Tuner tuners[4][4] = {
        {Tuner("-", {0}), Tuner("-", {0}), Tuner("-", {0}), Tuner("-", {0})},
        {Tuner("upfd_count_1_0", {0, 1}), Tuner("-", {0}), Tuner("-", {0}), Tuner("-", {0})},
        {Tuner("upfd_count_2_0", {0, 1}), Tuner("upfd_count_2_1", {0, 1}), Tuner("-", {0}), Tuner("-", {0})},
        {Tuner("upfd_count_3_0", {0, 1}), Tuner("upfd_count_3_1", {0, 1}), Tuner("upfd_count_3_2", {0, 1}),
         Tuner("-", {0})}};
    double bytes = (double) nuses * 8 + (double) nlows * 8;
    auto go = [&](auto kernel, hipEvent_t ea, hipEvent_t eb) {
      if (ea) {
        hipExtLaunchKernelGGL(kernel, dim3(nblocks(nuses)), dim3(BLOCK), 0, stream(), ea, eb, 0, nuses,
                              lows_of_highs, counts, slot.get());
        check_launch("upfd_count");
      } else {
        launch_timed("upfd_count", bytes, kernel, dim3(nblocks(nuses)), dim3(BLOCK), nuses, lows_of_highs, counts,
                     slot.get());
      }
    };
    tuned(tuners[hd][ld], nuses, [&](u32 v, hipEvent_t ea, hipEvent_t eb) {
      if (v) go(upfd_count<true>, ea, eb);
      else go(upfd_count<false>, ea, eb);
    });
  }
  u.offsets = exscan(counts, nlows);
  Buf<u32> tmp(nuses);
  if (nuses) {
    // down map and slots in, uses out, offsets read and counters cleared
    launch_timed("upfd_fill", (double) nuses * 12 + (double) nlows * 8, upfd_fill, dim3(nblocks(nuses)),
                 dim3(BLOCK), nuses, lows_of_highs, slot.get(), u.offsets.get(), counts, tmp.get());
    zero_counters().dirty = false;
  }
  slot.release();
  u.adj = Buf<u32>(nuses);
  u.dirs = Buf<u32>(nuses);
  if (nlows && nuses) {
    // offsets, uses in, (high, direction) out; lanes per bucket (4 to 32)
    // chosen online per (high, low) dimension pair and size class
    // variants: G lanes per bucket reading the fill buffer (4 to 32), or
    // (200 + S) the bucket held in S lanes (8 to 64)
    static std::vector<u32> const V = {4, 8, 16, 32, 208, 216, 232, 264};
    static thread_local Tuner tuners[4][4] = {
        {Tuner("-", {4}), Tuner("-", {4}), Tuner("-", {4}), Tuner("-", {4})},
        {Tuner("upfd_rank_split_1_0", V), Tuner("-", {4}), Tuner("-", {4}), Tuner("-", {4})},
        {Tuner("upfd_rank_split_2_0", V), Tuner("upfd_rank_split_2_1", V), Tuner("-", {4}), Tuner("-", {4})},
        {Tuner("upfd_rank_split_3_0", V), Tuner("upfd_rank_split_3_1", V), Tuner("upfd_rank_split_3_2", V),
         Tuner("-", {4})}};
    double bytes = (double) (nlows + 1) * 4 + (double) nuses * 12;
    u32 const* o = u.offsets.get();
    auto go = [&](auto kernel, u32 g, hipEvent_t ea, hipEvent_t eb) {
      dim3 grid(nblocks((size_t) nlows * g));
      if (ea) {
        hipExtLaunchKernelGGL(kernel, grid, dim3(BLOCK), 0, stream(), ea, eb, 0, nlows, deg, o, tmp.get(),
                              u.adj.get(), u.dirs.get());
        check_launch("upfd_rank_split");
      } else {
        launch_timed("upfd_rank_split", bytes, kernel, grid, dim3(BLOCK), nlows, deg, o, tmp.get(), u.adj.get(),
                     u.dirs.get());
      }
    };
    tuned(tuners[hd][ld], nlows, [&](u32 g, hipEvent_t ea, hipEvent_t eb) {
      if (g == 4) go(upfd_rank_split<4>, g, ea, eb);
      else if (g == 8) go(upfd_rank_split<8>, g, ea, eb);
      else if (g == 16) go(upfd_rank_split<16>, g, ea, eb);
      else if (g == 32) go(upfd_rank_split<32>, g, ea, eb);
      else if (g == 208) go(upfd_rank_seg<8>, 8, ea, eb);
      else if (g == 216) go(upfd_rank_seg<16>, 16, ea, eb);
      else if (g == 232) go(upfd_rank_seg<32>, 32, ea, eb);
      else go(upfd_rank_seg<64>, 64, ea, eb);
    });
  }
}

// ---------------------------------------------------------- reflect_down
// For every high entity and each of its canonical lows, the low with those
// vertices (reflect_down.cpp:13-81, find_by_verts.hpp:14-76).  The answer is
// unique in a conforming mesh, so any search finds the reference's: here an
// open-addressing hash table keyed by the sorted vertex tuple (lows inserted
// once, mostly one probe per lookup at the table's load) replaces the scan
// of the first vertex's upward adjacency, whose random reads grow with
// vertex valence.
template <u32 N>
__device__ inline void sort_small(u32 (&v)[N]) {
#pragma unroll
  for (u32 i = 1; i < N; ++i)
#pragma unroll
    for (u32 j = i; j > 0; --j)
      if (v[j - 1] > v[j]) {
        u32 t = v[j - 1];
        v[j - 1] = v[j];
        v[j] = t;
      }
}

template <u32 N>
__device__ inline u32 tuple_hash(u32 const (&v)[N], u32 mask) {
  unsigned long long h = 0x9E3779B97F4A7C15ull;
#pragma unroll
  for (u32 i = 0; i < N; ++i) {
    h ^= v[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ull;
  }
  h ^= h >> 31;
  return (u32) h & mask;
}

// Insertion, one launch (HOME_FIRST false): every low takes the first free
// slot from its home with compare-and-swap.  Or two launches (HOME_FIRST):
// first every low stores its id at its home slot with a plain store (where
// several share a home, one of them is left); then each low reads its home
// slot back and only the ones that lost take a free slot further on with
// compare-and-swap, as linear probing does.  A home slot is never empty
// after the first launch, so no loser can take it, and every low stays
// reachable from its home through occupied slots.  Most lows win at the
// table's load, so most atomics become plain stores and loads: faster on
// small and medium tables (64^3 x 6 box: 181 -> 151 us), slower on tables of 2^28
// slots and more, where the scattered plain stores cost more than the
// atomics; the tuner picks per size class.
template <u32 VPL>
__device__ inline u32 low_home(u32 e, u32 const* __restrict__ vl, u32 mask) {
  u32 v[VPL];
#pragma unroll
  for (u32 k = 0; k < VPL; ++k) v[k] = vl[(size_t) e * VPL + k];
  sort_small(v);
  return tuple_hash(v, mask);
}

template <u32 VPL>
__global__ __launch_bounds__(BLOCK) void lows_hash_home(u32 nl, u32 const* __restrict__ vl, u32 mask,
                                                        u32* __restrict__ table) {
  u32 e = xcd_block() * BLOCK + threadIdx.x;
  if (e >= nl) return;
  table[low_home<VPL>(e, vl, mask)] = e;
}

template <u32 VPL, bool HOME_FIRST>
__global__ __launch_bounds__(BLOCK) void lows_hash_insert(u32 nl, u32 const* __restrict__ vl, u32 mask,
                                                          u32* __restrict__ table, u32* __restrict__ err) {
  u32 e = xcd_block() * BLOCK + threadIdx.x;
  if (e >= nl) return;
  u32 slot = low_home<VPL>(e, vl, mask);
  if (HOME_FIRST) {
    if (table[slot] == e) return;
    slot = (slot + 1) & mask;
  }
  for (u32 probe = HOME_FIRST ? 1 : 0; probe <= mask; ++probe) {
    if (atomicCAS(&table[slot], INVALID, e) == INVALID) return;
    slot = (slot + 1) & mask;
  }
  *err = ERR_REFLECT_NOT_FOUND;
}

template <u32 HD, u32 LD>
__global__ __launch_bounds__(BLOCK) void reflect_down_hash(u32 nh, u32 const* __restrict__ vh,
                                                           u32 const* __restrict__ vl, u32 mask,
                                                           u32 const* __restrict__ table, u32* __restrict__ out,
                                                           u32* __restrict__ err) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= nh) return;
  constexpr u32 VPH = HD + 1, VPL = LD + 1;
  constexpr u32 LPH = (HD == 3) ? (LD == 1 ? 6 : 4) : 3;
  u32 hv[VPH];
#pragma unroll
  for (u32 k = 0; k < VPH; ++k) hv[k] = vh[(size_t) i * VPH + k];
  // the LPH lookups advance together one probe at a time, so each probe
  // level issues all its table reads, then all its vertex reads, before
  // waiting (two memory round trips per level instead of per lookup)
  u32 w[LPH][VPL], slot[LPH], found[LPH];
  bool open[LPH];
#pragma unroll
  for (u32 j = 0; j < LPH; ++j) {
    u32 const* lc = canon(HD, LD, 0, j);
#pragma unroll
    for (u32 k = 0; k < VPL; ++k) w[j][k] = hv[lc[k]];
    sort_small(w[j]);
    slot[j] = tuple_hash(w[j], mask);
    found[j] = INVALID;
    open[j] = true;
  }
  for (u32 probe = 0; probe <= mask; ++probe) {
    u32 e[LPH];
#pragma unroll
    for (u32 j = 0; j < LPH; ++j) e[j] = open[j] ? table[slot[j]] : INVALID;
    u32 t[LPH][VPL];
#pragma unroll
    for (u32 j = 0; j < LPH; ++j)
#pragma unroll
      for (u32 k = 0; k < VPL; ++k) t[j][k] = e[j] != INVALID ? vl[(size_t) e[j] * VPL + k] : INVALID;
    bool more = false;
#pragma unroll
    for (u32 j = 0; j < LPH; ++j) {
      if (!open[j]) continue;
      if (e[j] == INVALID) {  // an empty slot ends the search: not present
        open[j] = false;
        continue;
      }
      sort_small(t[j]);
      bool same = true;
#pragma unroll
      for (u32 k = 0; k < VPL; ++k) same = same && (t[j][k] == w[j][k]);
      if (same) {
        found[j] = e[j];
        open[j] = false;
      } else {
        slot[j] = (slot[j] + 1) & mask;
        more = true;
      }
    }
    if (!more) break;
  }
#pragma unroll
  for (u32 j = 0; j < LPH; ++j) {
    if (found[j] == INVALID) *err = ERR_REFLECT_NOT_FOUND;
    out[(size_t) i * LPH + j] = found[j];
  }
}

static Buf<u32> reflect_down(Mesh& m, u32 hd, u32 ld) {
  u32 nh = count(m, hd);
  u32 nl = count(m, ld);
  u32 const* vh = ask_down(m, hd, 0);
  u32 const* vl = ask_down(m, ld, 0);
  u32 lph = h_down_deg[hd][ld];
  Buf<u32> out((size_t) nh * lph);
  if (!nh) return out;
  u32* err = dev_error_word();
  // table of at least four times as many slots as lows (load factor <= 1/4:
  // most lookups end at their first probe); a 3D mesh's edge table serves
  // both reflections onto edges
#ifndef OHX_HASH_SLOTS_PER_LOW
#define OHX_HASH_SLOTS_PER_LOW 4
#endif
  if (!m.low_table[ld]) {
    OHX_REQUIRE((size_t) OHX_HASH_SLOTS_PER_LOW * nl <= ((size_t) 1 << 31),
                "reflect_down: too many lows for the table");
    u32 cap = 1024;
    while (cap < (size_t) OHX_HASH_SLOTS_PER_LOW * nl) cap *= 2;
    u32 mask = cap - 1;
    Buf<u32> table = filled(cap, INVALID);
    // compulsory bytes: low vertices in, the table written once
    double ins_bytes = (double) nl * (ld + 1) * 4 + (double) cap * 4;
    if (nl) {
      static thread_local Tuner tuners[3] = {Tuner("-", {0}), Tuner("lows_hash_insert_1", {0, 1}),
                                             Tuner("lows_hash_insert_2", {0, 1})};
      dim3 grid(nblocks(nl));
      u32* tp = table.get();
      // compulsory bytes of the home launch: low vertices in, home slots
      // written
      double home_bytes = (double) nl * (ld + 1) * 4 + (double) nl * 4;
      auto go = [&](auto home, auto insert, bool two, hipEvent_t ea, hipEvent_t eb) {
        if (ea) {
          if (two) {
            hipExtLaunchKernelGGL(home, grid, dim3(BLOCK), 0, stream(), ea, nullptr, 0, nl, vl, mask, tp);
            check_launch("lows_hash_home");
            hipExtLaunchKernelGGL(insert, grid, dim3(BLOCK), 0, stream(), nullptr, eb, 0, nl, vl, mask, tp, err);
          } else {
            hipExtLaunchKernelGGL(insert, grid, dim3(BLOCK), 0, stream(), ea, eb, 0, nl, vl, mask, tp, err);
          }
          check_launch("lows_hash_insert");
        } else {
          if (two) launch_timed("lows_hash_home", home_bytes, home, grid, dim3(BLOCK), nl, vl, mask, tp);
          launch_timed("lows_hash_insert", ins_bytes, insert, grid, dim3(BLOCK), nl, vl, mask, tp, err);
        }
      };
      tuned(tuners[ld], nl, [&](u32 v, hipEvent_t ea, hipEvent_t eb) {
        if (ld == 1) {
          if (v) go(lows_hash_home<2>, lows_hash_insert<2, true>, true, ea, eb);
          else go(lows_hash_home<2>, lows_hash_insert<2, false>, false, ea, eb);
        } else {
          if (v) go(lows_hash_home<3>, lows_hash_insert<3, true>, true, ea, eb);
          else go(lows_hash_home<3>, lows_hash_insert<3, false>, false, ea, eb);
        }
      });
    }
    m.low_table[ld] = std::move(table);
    m.low_mask[ld] = mask;
  }
  u32 const mask = m.low_mask[ld], cap = mask + 1;
  Buf<u32> const& table = m.low_table[ld];
  // compulsory bytes: high vertices in, low ids out, the table and the low
  // vertices read once
  double prof_bytes = (double) nh * ((hd + 1) * 4 + lph * 4) + (double) cap * 4 + (double) nl * (ld + 1) * 4;
  dim3 g(nblocks(nh)), b(BLOCK);
  if (hd == 3 && ld == 1)
    launch_timed("reflect_down", prof_bytes, reflect_down_hash<3, 1>, g, b, nh, vh, vl, mask, table.get(), out.get(),
                 err);
  else if (hd == 3 && ld == 2)
    launch_timed("reflect_down", prof_bytes, reflect_down_hash<3, 2>, g, b, nh, vh, vl, mask, table.get(), out.get(),
                 err);
  else if (hd == 2 && ld == 1)
    launch_timed("reflect_down", prof_bytes, reflect_down_hash<2, 1>, g, b, nh, vh, vl, mask, table.get(), out.get(),
                 err);
  else
    fail("reflect_down: unsupported dimensions", __FILE__, __LINE__);
  // only a 3D mesh's edges are found twice (from tets and from faces); any
  // other table goes back to the arena at once (stream-ordered)
  if (!(m.dim == 3 && ld == 1 && !m.down[hd == 3 ? 2 : 3][1])) {
    m.low_table[ld].release();
    m.low_mask[ld] = 0;
  }
  return out;
}

// the owning low entity of every use of an up adjacency (offsets over n
// lows): each low writes its index over its range
static Buf<u32> up_owner(u32 const* offs, u32 n, u32 nuses) {
  Buf<u32> own(nuses);
  u32* o = own.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    for (u32 j = offs[i]; j < offs[i + 1]; ++j) o[j] = i;
  });
  return own;
}

// ----------------------------------------------------------------- stars
// get_star_general (star.cpp:13-111): lows of the highs around each low,
// first occurrence kept, in up-adjacency order.  One wave per low entity:
// the lows of its highs are staged in LDS and lane p keeps entry p if it is
// not the entity itself and no earlier entry equals it; a ballot prefix
// gives each kept entry its place.  Entities with more than STAR_CAP
// entries (never in a simplicial mesh of sane valence) take a serial path.
constexpr u32 STAR_CAP = 1024;
constexpr u32 STAR_WAVES = BLOCK / 64;

template <bool FILL>
__global__ __launch_bounds__(BLOCK) void star_general_k(u32 nl, u32 const* __restrict__ uo, u32 const* __restrict__ ua,
                                                        u32 const* __restrict__ lh, u32 lph, u32* __restrict__ deg,
                                                        u32 const* __restrict__ go, u32* __restrict__ ga) {
  __shared__ u32 buf[STAR_WAVES][STAR_CAP];
  u32 w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  u32 i = blockIdx.x * STAR_WAVES + w;
  bool live = i < nl;
  u32 f = live ? uo[i] : 0, e = live ? uo[i + 1] : 0;
  u32 n = (e - f) * lph;
  bool staged = live && n <= STAR_CAP;
  u32* E = buf[w];
  if (staged)
    for (u32 p = lane; p < n; p += 64) E[p] = lh[(size_t) ua[f + p / lph] * lph + p % lph];
  __syncthreads();
  if (staged) {
    u32 cnt = 0;
    u32 base_out = FILL ? go[i] : 0;
    u64 below = (1ull << lane) - 1ull;
    for (u32 b = 0; b < n; b += 64) {
      u32 p = b + lane;
      bool keep = false;
      u32 s = 0;
      if (p < n) {
        s = E[p];
        keep = s != i;
        for (u32 q = 0; q < p && keep; ++q)
          if (E[q] == s) keep = false;
      }
      u64 m = __ballot(keep);
      if (FILL && keep) ga[base_out + cnt + (u32) __popcll(m & below)] = s;
      cnt += (u32) __popcll(m);
    }
    if (!FILL && lane == 0) deg[i] = cnt;
  } else if (live && lane == 0) {
    // serial: uniqueness by scanning what came before
    u32 cnt = 0;
    u32 base_out = FILL ? go[i] : 0;
    for (u32 p = 0; p < n; ++p) {
      u32 s = lh[(size_t) ua[f + p / lph] * lph + p % lph];
      if (s == i) continue;
      bool seen = false;
      for (u32 q = 0; q < p && !seen; ++q) seen = lh[(size_t) ua[f + q / lph] * lph + q % lph] == s;
      if (seen) continue;
      if (FILL) ga[base_out + cnt] = s;
      ++cnt;
    }
    if (!FILL) deg[i] = cnt;
  }
}

static Graph star_general(Mesh& m, u32 ld, u32 hd) {
  u32 nl = count(m, ld);
  Up const& u = ask_up(m, ld, hd);
  u32 const* lh = ask_down(m, hd, ld);
  u32 lph = h_down_deg[hd][ld];
  Buf<u32> deg(nl);
  u32 nb = (nl + STAR_WAVES - 1) / STAR_WAVES;
  Graph g;
  g.offsets = Buf<u32>(nl + 1);
  if (nl) {
    hipLaunchKernelGGL(star_general_k<false>, dim3(nb), dim3(BLOCK), 0, stream(), nl, u.offsets.get(), u.adj.get(), lh,
                       lph, deg.get(), (u32 const*) nullptr, (u32*) nullptr);
    check_launch("star_general");
  }
  u32 nadj = exscan_total(deg.get(), nl, g.offsets.get());
  g.adj = Buf<u32>(nadj);
  if (nl) {
    hipLaunchKernelGGL(star_general_k<true>, dim3(nb), dim3(BLOCK), 0, stream(), nl, u.offsets.get(), u.adj.get(), lh,
                       lph, (u32*) nullptr, g.offsets.get(), g.adj.get());
    check_launch("star_general");
  }
  return g;
}

// get_vertex_edge_star (star.cpp:113-153)
static Graph star_vertex_edge(Mesh& m) {
  u32 nv = count(m, 0);
  u32 const* ev = ask_down(m, 1, 0);
  Up const& u = ask_up(m, 0, 1);
  Graph g;
  g.offsets = copy(u.offsets.get(), nv + 1);
  u32 nadj = u.adj.size();
  g.adj = Buf<u32>(nadj);
  u32* ga = g.adj.get();
  u32 const* ua = u.adj.get();
  u32 const* ud = u.dirs.get();
  parallel_for(nadj, OHX_LAMBDA(u32 j) { ga[j] = ev[2 * (size_t) ua[j] + (1 - ud[j])]; });
  return g;
}

// get_edge_triangle_star (star.cpp:155-200)
static Graph star_edge_tri(Mesh& m) {
  u32 ne = count(m, 1);
  u32 const* et = ask_down(m, 2, 1);
  Up const& u = ask_up(m, 1, 2);
  Graph g;
  g.offsets = Buf<u32>(ne + 1);
  u32* go = g.offsets.get();
  u32 const* uo = u.offsets.get();
  parallel_for(ne + 1, OHX_LAMBDA(u32 i) { go[i] = uo[i] * 2; });
  u32 nuse = u.adj.size();
  g.adj = Buf<u32>((size_t) nuse * 2);
  u32* ga = g.adj.get();
  u32 const* ua = u.adj.get();
  u32 const* ud = u.dirs.get();
  parallel_for(nuse, OHX_LAMBDA(u32 j) {
    u32 t = ua[j], d = ud[j];
    ga[2 * j + 0] = et[3 * (size_t) t + (d + 1) % 3];
    ga[2 * j + 1] = et[3 * (size_t) t + (d + 2) % 3];
  });
  return g;
}

// get_edge_tet_star (star.cpp:202-278): the triangle star followed by the
// edge opposite to this one in every adjacent tet
static Graph star_edge_tet(Mesh& m) {
  u32 ne = count(m, 1);
  u32 const* et = ask_down(m, 2, 1);
  u32 const* er = ask_down(m, 3, 1);
  Up const& ut = ask_up(m, 1, 2);
  Up const& ur = ask_up(m, 1, 3);
  u32 const* to = ut.offsets.get();
  u32 const* ta = ut.adj.get();
  u32 const* td = ut.dirs.get();
  u32 const* ro = ur.offsets.get();
  u32 const* ra = ur.adj.get();
  u32 const* rd = ur.dirs.get();
  Buf<u32> deg(ne);
  u32* dg = deg.get();
  parallel_for(ne, OHX_LAMBDA(u32 i) { dg[i] = 2 * (to[i + 1] - to[i]) + (ro[i + 1] - ro[i]); });
  Graph g;
  g.offsets = Buf<u32>(ne + 1);
  size_t nadj = (size_t) 2 * ut.adj.size() + ur.adj.size();
  if (nadj >= ((size_t) 1 << 32)) fail("edge star: more than 2^32 entries", __FILE__, __LINE__);
  exscan(dg, ne, g.offsets.get());
  g.adj = Buf<u32>(nadj);
  u32* ga = g.adj.get();
  u32 const* go = g.offsets.get();
  // one thread per use: the owning edge's offsets come from the up lists'
  // own offsets (edge i's triangle uses fill go[i] + 2 (j - to[i]), its tet
  // uses follow them), so every read and write is coalesced
  u32 ntu = ut.adj.size(), nru = ur.adj.size();
  Buf<u32> owner_t = up_owner(to, ne, ntu), owner_r = up_owner(ro, ne, nru);
  u32 const* ot = owner_t.get();
  u32 const* orr = owner_r.get();
  parallel_for(ntu, OHX_LAMBDA(u32 j) {
    u32 i = ot[j], t = ta[j], d = td[j];
    u32 o = go[i] + 2 * (j - to[i]);
    ga[o] = et[3 * (size_t) t + (d + 1) % 3];
    ga[o + 1] = et[3 * (size_t) t + (d + 2) % 3];
  });
  parallel_for(nru, OHX_LAMBDA(u32 j) {
    u32 i = orr[j];
    u32 o = go[i] + 2 * (to[i + 1] - to[i]) + (j - ro[i]);
    ga[o] = er[6 * (size_t) ra[j] + opp(3, 1)[rd[j]]];
  });
  return g;
}

// mesh_get_star (star.cpp:280-315)
static Graph get_star(Mesh& m, u32 ld, u32 hd) {
  if (ld == 0 && hd == m.dim && m.dim != 1 && has_dim(m, 1)) return get_star(m, 0, 1);
  if (ld == 0 && hd == 1) return star_vertex_edge(m);
  if (ld == 1 && hd == 2) return star_edge_tri(m);
  if (ld == 1 && hd == 3 && has_dim(m, 2)) return star_edge_tet(m);
  return star_general(m, ld, hd);
}

// ------------------------------------------------------------------ dual
// element_dual_from_verts (dual.cpp:40-98): across side j of element i,
// the element of the side's first vertex (other than i, first in up order)
// that also contains every other side vertex.  The up lists are sorted by
// element, so this is an intersection of sorted lists, walked in the first
// vertex's order.
static Buf<u32> dual_from_verts(Mesh& m) {
  u32 dim = m.dim, sd = dim - 1;
  u32 ne = count(m, dim);
  u32 const* ev = ask_down(m, dim, 0);
  Up const& u = ask_up(m, 0, dim);
  u32 vpe = dim + 1, spe = h_down_deg[dim][sd], vps = h_down_deg[sd][0];
  Buf<u32> out((size_t) ne * spe);
  u32* o = out.get();
  u32 const* uo = u.offsets.get();
  u32 const* ua = u.adj.get();
  u32* err = dev_error_word();
  parallel_for(ne, OHX_LAMBDA(u32 i) {
    for (u32 j = 0; j < spe; ++j) {
      u32 const* sc = canon(dim, sd, 0, j);
      u32 pos[3], end[3];
      for (u32 k = 0; k < vps; ++k) {
        u32 v = ev[(size_t) i * vpe + sc[k]];
        pos[k] = uo[v];
        end[k] = uo[v + 1];
      }
      u32 found = INVALID, nfound = 0;
      bool done = false;
      for (u32 a = pos[0]; a < end[0] && !done; ++a) {
        u32 c = ua[a];
        bool all = true;
        for (u32 k = 1; k < vps; ++k) {
          while (pos[k] < end[k] && ua[pos[k]] < c) ++pos[k];
          if (pos[k] == end[k]) {
            done = true;
            all = false;
            break;
          }
          if (ua[pos[k]] != c) all = false;
        }
        if (all && c != i) {
          if (!nfound) found = c;
          ++nfound;
        }
      }
      if (nfound > 1) *err = ERR_DUAL_NONMANIFOLD;
      o[(size_t) i * spe + j] = found;
    }
  });
  return out;
}

// element_dual_from_sides (dual.cpp:116-156)
static Buf<u32> dual_from_sides(Mesh& m) {
  u32 dim = m.dim;
  u32 ne = count(m, dim);
  u32 spe = h_down_deg[dim][dim - 1];
  u32 const* se = ask_down(m, dim, dim - 1);
  Up const& u = ask_up(m, dim - 1, dim);
  Buf<u32> out((size_t) ne * spe);
  u32* o = out.get();
  u32 const* uo = u.offsets.get();
  u32 const* ua = u.adj.get();
  parallel_for(ne, OHX_LAMBDA(u32 i) {
    for (u32 j = 0; j < spe; ++j) {
      u32 s = se[(size_t) i * spe + j];
      u32 r = INVALID;
      for (u32 k = uo[s]; k < uo[s + 1]; ++k)
        if (ua[k] != i) {
          r = ua[k];
          break;
        }
      o[(size_t) i * spe + j] = r;
    }
  });
  return out;
}

// The same dual, unique in a conforming mesh (at most two elements per
// side), by pairing the elements' sides in a hash table keyed by the sorted
// side vertices: the first of a pair takes a slot, the second finds it and
// writes both entries; sides without a partner stay INVALID.  It needs
// neither the sides' numbering (down(dim, dim - 1)) nor their upward map,
// which a pass that only wants the dual would otherwise build.
template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void dual_pairs_k(u32 nsides, u32 const* __restrict__ ev, u32 mask,
                                                      u32* __restrict__ table, u32* __restrict__ dual) {
  u32 t = xcd_block() * BLOCK + threadIdx.x;
  if (t >= nsides) return;
  constexpr u32 VPE = DIM + 1, SPE = DIM + 1;
  u32 const i = t / SPE, j = t % SPE;
  u32 const* sc = canon(DIM, DIM - 1, 0, j);
  u32 key[DIM];
#pragma unroll
  for (u32 k = 0; k < DIM; ++k) key[k] = ev[(size_t) i * VPE + sc[k]];
  sort_small(key);
  u32 h = tuple_hash(key, mask);
  for (u32 probe = 0; probe <= mask; ++probe) {
    u32 prev = atomicCAS(&table[h], INVALID, t);
    if (prev == INVALID) return;  // first of its pair: the partner writes both
    u32 pi = prev / SPE;
    u32 const* pc = canon(DIM, DIM - 1, 0, prev % SPE);
    u32 pk[DIM];
#pragma unroll
    for (u32 k = 0; k < DIM; ++k) pk[k] = ev[(size_t) pi * VPE + pc[k]];
    sort_small(pk);
    bool same = true;
#pragma unroll
    for (u32 k = 0; k < DIM; ++k) same = same && pk[k] == key[k];
    if (same) {
      dual[t] = pi;
      dual[prev] = i;
      return;
    }
    h = (h + 1) & mask;
  }
}

// The same dual in 3D from the tets around an edge of each face: the tet
// across face k of tet t is the other tet around the face's first edge
// that holds the face's third vertex.  Used when up(1,3) and down(3,1)
// are there already (a swap pass builds them): no table, no atomics.
__global__ __launch_bounds__(BLOCK) void dual_edges_k(u32 nt, u32 const* __restrict__ tv, u32 const* __restrict__ te,
                                                      u32 const* __restrict__ uo, u32 const* __restrict__ ua,
                                                      u32* __restrict__ dual) {
  u32 s = xcd_block() * BLOCK + threadIdx.x;
  if (s >= 4 * nt) return;
  u32 const t = s / 4, k = s % 4;
  u32 const* fv = canon(3, 2, 0, k);
  // the tet's local edge between the face's first two vertices
  u32 le = 0;
#pragma unroll
  for (u32 l = 0; l < 6; ++l) {
    u32 const* lv = canon(3, 1, 0, l);
    if ((lv[0] == fv[0] && lv[1] == fv[1]) || (lv[0] == fv[1] && lv[1] == fv[0])) le = l;
  }
  u32 const edge = te[6 * (size_t) t + le];
  u32 const c = tv[4 * (size_t) t + fv[2]];
  u32 r = INVALID;
  for (u32 j = uo[edge]; j < uo[edge + 1]; ++j) {
    u32 o = ua[j];
    if (o == t) continue;
    u32 const* ov = tv + 4 * (size_t) o;
    if (ov[0] == c || ov[1] == c || ov[2] == c || ov[3] == c) {
      r = o;
      break;
    }
  }
  dual[s] = r;
}

static Buf<u32> dual_by_pairs(Mesh& m) {
  u32 dim = m.dim;
  u32 ne = count(m, dim);
  u32 spe = h_down_deg[dim][dim - 1];
  size_t nsides = (size_t) ne * spe;
  Buf<u32> out(nsides);
  if (!nsides) return out;
  OHX_REQUIRE(nsides < ((size_t) 1 << 31), "dual: too many element sides");
  u32 cap = 1;
  while (cap < 2 * nsides) cap <<= 1;
  Buf<u32> table(cap);
  fill(table.get(), cap, INVALID);
  fill(out.get(), (u32) nsides, INVALID);
  u32 const* ev = ask_down(m, dim, 0);
  if (dim == 3)
    hipLaunchKernelGGL(dual_pairs_k<3>, dim3(nblocks(nsides)), dim3(BLOCK), 0, stream(), (u32) nsides, ev, cap - 1,
                       table.get(), out.get());
  else
    hipLaunchKernelGGL(dual_pairs_k<2>, dim3(nblocks(nsides)), dim3(BLOCK), 0, stream(), (u32) nsides, ev, cap - 1,
                       table.get(), out.get());
  check_launch("dual_pairs");
  return out;
}

// bridge_graph_general (bridge_graph.cpp:12-80)
void bridge_graph(u32 n, u32 const* offs, u32 const* adj, u32* p_nedges, Buf<u32>* ev, Buf<u32>* dirs) {
  Buf<u32> deg(n);
  u32* dg = deg.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    u32 c = 0;
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      if (i < adj[j]) ++c;
    dg[i] = c;
  });
  Buf<u32> bo(n + 1);
  u32 ne = exscan_total(dg, n, bo.get());
  *ev = Buf<u32>((size_t) ne * 2);
  if (dirs) *dirs = Buf<u32>(ne);
  u32* e2 = ev->get();
  u32* ed = dirs ? dirs->get() : nullptr;
  u32 const* b = bo.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    u32 e = b[i];
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      if (i < adj[j]) {
        e2[2 * (size_t) e + 0] = i;
        e2[2 * (size_t) e + 1] = adj[j];
        if (ed) ed[e] = j - offs[i];
        ++e;
      }
  });
  *p_nedges = ne;
}

// ------------------------------------------------------------ lazy caches
// OHX_TRACE_DERIVE=1 prints every derivation (kind, dimensions, size) to
// stderr: which adjacencies each pass rebuilds
static bool trace_derive() {
  static int on = -1;
  if (on < 0) {
    char const* e = std::getenv("OHX_TRACE_DERIVE");
    on = (e && e[0] == '1') ? 1 : 0;
  }
  return on == 1;
}

// mesh_ask_down (mesh.cpp:136-181)
u32 const* ask_down(Mesh& m, u32 h, u32 l) {
  OHX_REQUIRE(h <= m.dim && l <= h, "ask_down: bad dimensions");
  if (m.down[h][l]) return m.down[h][l].get();
  if (trace_derive()) std::fprintf(stderr, "derive down %u %u n=%u\n", h, l, m.counts[h]);
  if (l == h) {
    m.down[h][l] = iota(m.counts[h]);
  } else if (l > 0) {
    m.down[h][l] = reflect_down(m, h, l);
  } else {
    OHX_REQUIRE(!m.full, "ask_down: missing entities in a full mesh");
    if (h == 1 && m.dim == 3) {
      // 3D edges from the vertex-to-vertex star across tets
      Graph const& g = ask_star(m, 0, m.dim);
      u32 ne;
      Buf<u32> ev;
      bridge_graph(m.counts[0], g.offsets.get(), g.adj.get(), &ne, &ev, nullptr);
      set_ents(m, 1, ne, std::move(ev));
    } else {
      // sides: bridge the element dual, then take each side's vertices
      // from its first element (bridge_graph.cpp:82-108, derive_sides.cpp:10-52)
      OHX_REQUIRE(h == m.dim - 1, "ask_down: cannot derive this dimension");
      u32 const* dual = ask_dual(m);
      u32 dim = m.dim, ne = m.counts[dim];
      u32 spe = h_down_deg[dim][dim - 1];
      Buf<u32> offs(ne + 1);
      u32* op = offs.get();
      parallel_for(ne + 1, OHX_LAMBDA(u32 i) { op[i] = i * spe; });
      u32 ns;
      Buf<u32> es, ess;
      bridge_graph(ne, op, dual, &ns, &es, &ess);
      u32 const* ev = m.down[dim][0].get();
      u32 vps = h_down_deg[dim - 1][0], vpe = dim + 1;
      Buf<u32> sv((size_t) ns * vps);
      u32* s = sv.get();
      u32 const* e = es.get();
      u32 const* ed = ess.get();
      parallel_for(ns, OHX_LAMBDA(u32 i) {
        u32 el = e[2 * (size_t) i];
        u32 const* sc = canon(dim, dim - 1, 0, ed[i]);
        for (u32 j = 0; j < vps; ++j) s[(size_t) i * vps + j] = ev[(size_t) el * vpe + sc[j]];
      });
      set_ents(m, h, ns, std::move(sv));
    }
  }
  return m.down[h][l].get();
}

// mesh_ask_up (mesh.cpp:190-217)
Up const& ask_up(Mesh& m, u32 l, u32 h) {
  if (m.up[l][h]) return *m.up[l][h];
  if (trace_derive()) std::fprintf(stderr, "derive up %u %u n=%u\n", l, h, m.counts[h]);
  auto u = std::make_unique<Up>();
  if (l == h) {
    u32 n = m.counts[h];
    u->offsets = iota(n + 1);
    u->adj = iota(n);
    u->dirs = filled(n, 0);
  } else {
    u32 const* lh = ask_down(m, h, l);
    up_from_down(h, l, m.counts[h], m.counts[l], lh, *u);
  }
  m.up[l][h] = std::move(u);
  return *m.up[l][h];
}

// mesh_ask_star (mesh.cpp:226-244)
Graph const& ask_star(Mesh& m, u32 l, u32 h) {
  if (m.star[l][h]) return *m.star[l][h];
  if (trace_derive()) std::fprintf(stderr, "derive star %u %u n=%u\n", l, h, m.counts[l]);
  auto g = std::make_unique<Graph>();
  if (l == h) {
    g->offsets = filled(m.counts[l] + 1, 0);
    g->adj = Buf<u32>(1);
  } else {
    *g = get_star(m, l, h);
  }
  m.star[l][h] = std::move(g);
  return *m.star[l][h];
}

// mesh_ask_dual (mesh.cpp:252-261)
u32 const* ask_dual(Mesh& m) {
  if (!m.dual) {
    u32 d = m.dim;
    if (!has_dim(m, d - 1))
      m.dual = dual_from_verts(m);
    else if (m.down[d][d - 1] && m.up[d - 1][d])
      m.dual = dual_from_sides(m);  // the side maps are there already
    else if (d == 3 && m.down[3][1] && m.up[1][3]) {
      u32 nt = m.counts[3];
      Buf<u32> out((size_t) 4 * nt);
      if (nt) {
        Up const& u = *m.up[1][3];
        hipLaunchKernelGGL(dual_edges_k, dim3(nblocks((size_t) 4 * nt)), dim3(BLOCK), 0, stream(), nt,
                           m.down[3][0].get(), m.down[3][1].get(), u.offsets.get(), u.adj.get(), out.get());
        check_launch("dual_edges");
      }
      m.dual = std::move(out);
    } else
      m.dual = dual_by_pairs(m);
  }
  return m.dual.get();
}

// ----------------------------------------------------------------- marks
Buf<u32> mark_down(Mesh& m, u32 hd, u32 ld, u32 const* mh) {
  u32 nl = count(m, ld);
  Up const& u = ask_up(m, ld, hd);
  Buf<u32> out(nl);
  u32* o = out.get();
  u32 const* uo = u.offsets.get();
  u32 const* ua = u.adj.get();
  parallel_for(nl, OHX_LAMBDA(u32 i) {
    u32 r = 0;
    for (u32 j = uo[i]; j < uo[i + 1]; ++j)
      if (mh[ua[j]]) {
        r = 1;
        break;
      }
    o[i] = r;
  });
  return out;
}

Buf<u32> mark_up(Mesh& m, u32 ld, u32 hd, u32 const* ml) {
  u32 nh = count(m, hd);
  u32 lph = h_down_deg[hd][ld];
  u32 const* lh = ask_down(m, hd, ld);
  Buf<u32> out(nh);
  u32* o = out.get();
  parallel_for(nh, OHX_LAMBDA(u32 i) {
    u32 r = 0;
    for (u32 j = 0; j < lph; ++j)
      if (ml[lh[(size_t) i * lph + j]]) {
        r = 1;
        break;
      }
    o[i] = r;
  });
  return out;
}

// mesh_mark_dual_layers (mark.cpp:90-126)
void mark_dual_layers(Mesh& m, Buf<u32>& marked, u32 nlayers) {
  u32 dim = m.dim;
  u32 ne = count(m, dim);
  u32 const* dual = ask_dual(m);
  u32 deg = h_down_deg[dim][dim - 1];
  for (u32 l = 0; l < nlayers; ++l) {
    Buf<u32> out(ne);
    u32* o = out.get();
    u32 const* in = marked.get();
    parallel_for(ne, OHX_LAMBDA(u32 i) {
      u32 r = in[i];
      for (u32 j = 0; j < deg; ++j) {
        u32 e = dual[(size_t) i * deg + j];
        if (e != INVALID && in[e]) r = 1;
      }
      o[i] = r;
    });
    marked = std::move(out);
  }
}

Buf<u32> mark_class(Mesh& m, u32 ent_dim, u32 target_dim, u32 target_id) {
  u32 n = count(m, ent_dim);
  u32 const* cd = u32_tag(m, ent_dim, "class_dim");
  u32 const* ci = (target_id != INVALID) ? u32_tag(m, ent_dim, "class_id") : nullptr;
  Buf<u32> out(n);
  u32* o = out.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    u32 r = 0;
    if (cd[i] == target_dim)
      if (target_id == INVALID || ci[i] == target_id) r = 1;
    o[i] = r;
  });
  return out;
}

void unmark_boundary(Mesh& m, u32 ent_dim, u32* marked) {
  u32 const* cd = u32_tag(m, ent_dim, "class_dim");
  u32 dim = m.dim;
  parallel_for(count(m, ent_dim), OHX_LAMBDA(u32 i) {
    if (cd[i] != dim) marked[i] = 0;
  });
}

// mesh_mark_slivers (mark.cpp:202-232): quality fused with the threshold
template <u32 DIM>
__global__ __launch_bounds__(BLOCK) void mark_sliver_f64(u32 n, u32 const* __restrict__ ev,
                                                         double const* __restrict__ x,
                                                         double const* __restrict__ mt, double good,
                                                         u32* __restrict__ out) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  out[i] = (elem_q<DIM>(i, ev, x, mt) < good) ? 1u : 0u;
}

Buf<u32> mark_slivers(Mesh& m, double good_qual, u32 nlayers, double const* eq, double const* gate) {
  u32 ne = count(m, m.dim);
  Buf<u32> s(ne);
  if (ne && eq) {  // the qualities are given: the threshold alone
    u32* o = s.get();
    parallel_for(ne, OHX_LAMBDA(u32 i) { o[i] = (eq[i] < good_qual && (!gate || *gate < good_qual)) ? 1u : 0u; });
  } else if (ne) {
    double prof_bytes = (double) ne * (m.dim + 1) * 4 + (double) count(m, 0) * 24 + (double) ne * 4;
    if (m.dim == 3)
      launch_timed("mark_sliver_f64", prof_bytes, mark_sliver_f64<3>, dim3(nblocks(ne)), dim3(BLOCK), ne,
                         ask_down(m, 3, 0), coords(m), metric_of(m), good_qual, s.get());
    else
      launch_timed("mark_sliver_f64", prof_bytes, mark_sliver_f64<2>, dim3(nblocks(ne)), dim3(BLOCK), ne,
                         ask_down(m, 2, 0), coords(m), metric_of(m), good_qual, s.get());
  }
  mark_dual_layers(m, s, nlayers);
  return s;
}

Buf<u32> mark_part_boundary(Mesh& m) {
  u32 dim = m.dim;
  u32 ns = count(m, dim - 1);
  Up const& u = ask_up(m, dim - 1, dim);
  u32 const* uo = u.offsets.get();
  Buf<u32> out(ns);
  u32* o = out.get();
  parallel_for(ns, OHX_LAMBDA(u32 i) { o[i] = (uo[i + 1] - uo[i]) < 2 ? 1u : 0u; });
  return out;
}

// ------------------------------------------------------ classification
// mesh_derive_model (derive_model.cpp:24-448).  Setup, run once per mesh.
// Normals and hinge dot products are computed on the device; the crease
// test acos(d) > angle and the connected-component BFS run on the host with
// the C library, as the reference runs its BFS on the host (bfs.cpp).
static void derive_class_dim(Mesh& m, double crease_angle) {
  u32 dim = m.dim;
  u32 ne = count(m, dim);
  add_tag_u32(m, dim, "class_dim", 1, filled(ne, dim));
  Buf<u32> bs = mark_part_boundary(m);
  u32 ns = count(m, dim - 1);
  Buf<u32> scd(ns);
  {
    u32* s = scd.get();
    u32 const* b = bs.get();
    parallel_for(ns, OHX_LAMBDA(u32 i) { s[i] = dim - b[i]; });
  }
  add_tag_u32(m, dim - 1, "class_dim", 1, std::move(scd));
  double const* x = coords(m);
  u32 const* sv = ask_down(m, dim - 1, 0);
  Buf<double> nrm((size_t) 3 * ns);
  {
    double* n = nrm.get();
    u32 const* b = bs.get();
    parallel_for(ns, OHX_LAMBDA(u32 i) {
      if (!b[i]) return;
      V3 v;
      if (dim == 2) {
        v = sub(ld3(x + 3 * (size_t) sv[2 * i + 1]), ld3(x + 3 * (size_t) sv[2 * i]));
      } else {
        V3 a = ld3(x + 3 * (size_t) sv[3 * i]);
        v = cross(sub(ld3(x + 3 * (size_t) sv[3 * i + 1]), a), sub(ld3(x + 3 * (size_t) sv[3 * i + 2]), a));
      }
      double s = 1.0 / sqrt(dot(v, v));  // normalize_vector (algebra.hpp:133-139)
      n[3 * (size_t) i + 0] = v.x * s;
      n[3 * (size_t) i + 1] = v.y * s;
      n[3 * (size_t) i + 2] = v.z * s;
    });
  }
  Buf<u32> bh = mark_down(m, dim - 1, dim - 2, bs.get());
  u32 nh = count(m, dim - 2);
  Up const& hs = ask_up(m, dim - 2, dim - 1);
  Buf<double> dots(nh);
  {
    double* d = dots.get();
    double const* n = nrm.get();
    u32 const* b = bs.get();
    u32 const* h = bh.get();
    u32 const* ho = hs.offsets.get();
    u32 const* ha = hs.adj.get();
    u32* err = dev_error_word();
    parallel_for(nh, OHX_LAMBDA(u32 i) {
      if (!h[i]) {
        d[i] = 2.0;  // not a boundary hinge: no angle
        return;
      }
      u32 nadj = 0, adj[2] = {0, 0};
      for (u32 j = ho[i]; j < ho[i + 1]; ++j)
        if (b[ha[j]]) {
          if (nadj < 2) adj[nadj] = ha[j];
          else *err = ERR_HINGE_SIDES;
          ++nadj;
        }
      d[i] = dot(ld3(n + 3 * (size_t) adj[0]), ld3(n + 3 * (size_t) adj[1]));
    });
  }
  std::vector<double> hd = host_copy(dots.get(), nh);
  std::vector<u32> crease(nh);
  for (u32 i = 0; i < nh; ++i) crease[i] = (hd[i] != 2.0) ? (std::acos(hd[i]) > crease_angle) : 0u;
  Buf<u32> cr = to_device(crease.data(), nh);
  Buf<u32> hcd(nh);
  {
    u32* o = hcd.get();
    u32 const* h = bh.get();
    u32 const* c = cr.get();
    parallel_for(nh, OHX_LAMBDA(u32 i) { o[i] = dim - h[i] - c[i]; });
  }
  if (dim == 2) {
    add_tag_u32(m, 0, "class_dim", 1, std::move(hcd));
    return;
  }
  u32 nv = count(m, 0);
  Up const& ve = ask_up(m, 0, 1);
  Buf<u32> vcd(nv);
  {
    u32* o = vcd.get();
    u32 const* vo = ve.offsets.get();
    u32 const* va = ve.adj.get();
    u32 const* c = cr.get();
    u32 const* hc = hcd.get();
    parallel_for(nv, OHX_LAMBDA(u32 i) {
      u32 nin = 0, d = 3;
      for (u32 j = vo[i]; j < vo[i + 1]; ++j) {
        u32 e = va[j];
        if (c[e]) ++nin;
        if (hc[e] < d) d = hc[e];
      }
      o[i] = d - (nin > 2 ? 1u : 0u);
    });
  }
  add_tag_u32(m, 1, "class_dim", 1, std::move(hcd));
  add_tag_u32(m, 0, "class_dim", 1, std::move(vcd));
}

// connected components numbered by their lowest node (bfs.cpp:18-103)
static void host_components(u32 n, std::vector<u32> const& offs, std::vector<u32> const& adj,
                            std::vector<u32>& comp) {
  std::vector<u32> layer(n, INVALID), queue(n);
  comp.assign(n, 0);
  u32 begin = 0, end = 0, c = 0;
  for (u32 i = 0; i < n; ++i) {
    if (layer[i] != INVALID) continue;
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
}

// set_equal_order_class_id (derive_model.cpp:270-396)
static void equal_order_class_id(Mesh& m, u32 d) {
  u32 n = count(m, d);
  Buf<u32> eq = mark_class(m, d, d, INVALID);
  Buf<u32> eqo(n + 1);
  u32 neq = exscan_total(eq.get(), n, eqo.get());
  Buf<u32> br = mark_class(m, d - 1, d, INVALID);
  u32 bpe = h_down_deg[d][d - 1];
  u32 const* be = ask_down(m, d, d - 1);
  Up const& u = ask_up(m, d - 1, d);
  Buf<u32> deg = filled(neq, 0);
  u32 const* e = eqo.get();
  u32 const* b = br.get();
  {
    u32* dg = deg.get();
    parallel_for(n, OHX_LAMBDA(u32 i) {
      if (e[i] == e[i + 1]) return;
      u32 k = 0;
      for (u32 j = 0; j < bpe; ++j)
        if (b[be[(size_t) i * bpe + j]]) ++k;
      dg[e[i]] = k;
    });
  }
  Buf<u32> offs(neq + 1);
  u32 nadj = exscan_total(deg.get(), neq, offs.get());
  Buf<u32> adj(nadj);
  {
    u32* a = adj.get();
    u32 const* o = offs.get();
    u32 const* uo = u.offsets.get();
    u32 const* ua = u.adj.get();
    parallel_for(n, OHX_LAMBDA(u32 i) {
      if (e[i] == e[i + 1]) return;
      u32 k = o[e[i]];
      for (u32 j = 0; j < bpe; ++j) {
        u32 s = be[(size_t) i * bpe + j];
        if (!b[s]) continue;
        for (u32 l = uo[s]; l < uo[s + 1]; ++l) {
          u32 other = ua[l];
          if (other == i) continue;
          if (e[other] != e[other + 1]) {
            a[k++] = e[other];
            break;
          }
        }
      }
    });
  }
  std::vector<u32> ho = host_copy(offs.get(), neq + 1), ha = host_copy(adj.get(), nadj), comp;
  host_components(neq, ho, ha, comp);
  Buf<u32> dc = to_device(comp.data(), neq);
  Buf<u32> cid = filled(n, 0);
  {
    u32* c = cid.get();
    u32 const* cc = dc.get();
    parallel_for(n, OHX_LAMBDA(u32 i) {
      if (e[i] != e[i + 1]) c[i] = cc[e[i]];
    });
  }
  add_tag_u32(m, d, "class_id", 1, std::move(cid));
}

// project_class_id_onto (derive_model.cpp:398-416): last matching high wins
static void project_class_id(Mesh& m, u32 ld) {
  u32 nl = count(m, ld);
  Up const& u = ask_up(m, ld, ld + 1);
  u32 const* hcd = u32_tag(m, ld + 1, "class_dim");
  u32 const* hci = u32_tag(m, ld + 1, "class_id");
  u32 const* lcd = u32_tag(m, ld, "class_dim");
  u32* lci = find_tag(m, ld, "class_id")->u.get();
  u32 const* uo = u.offsets.get();
  u32 const* ua = u.adj.get();
  parallel_for(nl, OHX_LAMBDA(u32 i) {
    for (u32 j = uo[i]; j < uo[i + 1]; ++j) {
      u32 h = ua[j];
      if (hcd[h] == lcd[i]) lci[i] = hci[h];
    }
  });
}

void derive_model(Mesh& m, double crease_angle) {
  derive_class_dim(m, crease_angle);
  u32 nv = count(m, 0);
  u32 const* vcd = u32_tag(m, 0, "class_dim");
  Buf<u32> is0(nv);
  {
    u32* z = is0.get();
    parallel_for(nv, OHX_LAMBDA(u32 i) { z[i] = vcd[i] == 0 ? 1u : 0u; });
  }
  Buf<u32> sc = exscan(is0.get(), nv);
  Buf<u32> vci(nv);
  {
    u32* c = vci.get();
    u32 const* z = is0.get();
    u32 const* s = sc.get();
    parallel_for(nv, OHX_LAMBDA(u32 i) { c[i] = z[i] ? s[i] : 0u; });
  }
  add_tag_u32(m, 0, "class_id", 1, std::move(vci));
  for (u32 d = 1; d <= m.dim; ++d) equal_order_class_id(m, d);
  for (u32 dd = 1; dd <= m.dim; ++dd) project_class_id(m, m.dim - dd);
}

}  // namespace ohx
