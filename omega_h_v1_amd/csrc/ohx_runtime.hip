// Runtime pieces: stream, pinned staging, caching HBM arena, event timers.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "ohx_common.hpp"

#include <algorithm>
#include <chrono>

namespace ohx {

static thread_local std::string g_last_error;
thread_local hipStream_t g_stream_override = nullptr;

void fail(char const* what, char const* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "ohx: %s (%s:%d)", what, file, line);
  g_last_error = buf;
  throw Error(buf);
}

std::string const& last_error() { return g_last_error; }
void set_last_error(std::string const& s) { g_last_error = s; }

// threads that drive meshes live until the process ends (the worker pool
// never retires them), so a thread's stream and staging are never released
Runtime& rt() {
  static thread_local Runtime r;
  if (!r.stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      throw Error("ohx: no HIP device visible (this library has no CPU path)");
    OHX_CHECK(hipGetDevice(&r.device));
    OHX_CHECK(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
    OHX_CHECK(hipHostMalloc(&r.pinned, 4096, hipHostMallocMapped));
    std::memset(r.pinned, 0, 4096);
    OHX_CHECK(hipHostGetDevicePointer(&r.pinned_dev, r.pinned, 0));
  }
  return r;
}

static char const* dev_error_text(unsigned e) {
  switch (e) {
    case ERR_REFLECT_NOT_FOUND: return "reflect_down: a sub-entity was not found (non-conforming mesh)";
    case ERR_DUAL_NONMANIFOLD: return "dual_from_verts: non-manifold side";
    case ERR_HINGE_SIDES: return "derive_model: more than two boundary sides at a hinge";
    case ERR_REFINE_QUALITY: return "refine: a child element has non-positive quality (refine_qualities.cpp:72)";
    case ERR_EXPOSED_SIDE: return "check_for_exposed_side: edge vertex not in entity";
    case ERR_NO_DESTINATION: return "vertex_dest: no collapse destination (collapses_to_verts.cpp:100)";
    case ERR_EDGE_RING: return "find_edge_ring: ring does not close or has fewer than 3 tets";
    case ERR_SCAN_SPIN: return "scan: look-back did not complete (spin limit)";
  }
  return "unknown device error";
}

#ifdef OHX_TRACE_SYNC
}  // namespace ohx
#include <execinfo.h>
namespace ohx {
#endif

void sync() {
#ifdef OHX_TRACE_SYNC
  if (prof_on()) {
    void* fr[6];
    int k = backtrace(fr, 6);
    backtrace_symbols_fd(fr + 1, k - 1, 2);
    std::fprintf(stderr, "--\n");
  }
#endif
  if (prof_on()) {
    // host-side count and wall time of the stream synchronisations
    auto t0 = std::chrono::steady_clock::now();
    OHX_CHECK(hipStreamSynchronize(stream()));
    auto t1 = std::chrono::steady_clock::now();
    prof_host("host_sync", std::chrono::duration<double, std::milli>(t1 - t0).count());
  } else {
    OHX_CHECK(hipStreamSynchronize(stream()));
  }
  volatile unsigned* e = host_slot<unsigned>(SLOT_ERROR);
  if (*e) {
    unsigned code = *e;
    *e = 0;
    fail(dev_error_text(code), __FILE__, __LINE__);
  }
}

void check_launch(char const* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(name, __FILE__, __LINE__);
}

namespace {
struct Ptrs16 {
  unsigned const* p[16];
};
__global__ void gather_scalars(Ptrs16 q, unsigned cnt, unsigned* o) {
  unsigned i = threadIdx.x;
  if (i < cnt) o[2 * i] = q.p[i] ? *q.p[i] : 0u;
}
}  // namespace

void read_many(unsigned const* const* ptrs, unsigned n, unsigned* out) {
  OHX_REQUIRE(n <= 16, "read_many: at most 16 scalars");
  if (!n) return;
  Ptrs16 pk;
  for (unsigned i = 0; i < 16; ++i) pk.p[i] = i < n ? ptrs[i] : nullptr;
  hipLaunchKernelGGL(gather_scalars, dim3(1), dim3(64), 0, stream(), pk, n, dev_slot<unsigned>(SLOT_GATHER));
  check_launch("read_many");
  sync();
  unsigned* h = host_slot<unsigned>(SLOT_GATHER);
  for (unsigned i = 0; i < n; ++i) out[i] = h[2 * i];
}

// ------------------------------------------------------------------ arena
// One caching arena for the process, guarded by a mutex.  A freed block is
// reused only by allocations on the stream it was freed on: work on that
// stream is ordered after the block's last use, so no synchronisation is
// needed, and streams of different threads never hand blocks to each other
// while their kernels may still read them.
namespace {
// Blocks come in power-of-two classes (256 B and up), indexed by log2.  The
// free blocks are kept per stream and class in plain vectors (a stream's
// entry found by a short linear search: one per driving thread), and the
// live blocks' classes in an open-addressing table keyed by address, so an
// allocation or a release costs a few array operations under the lock.
constexpr unsigned ARENA_CLASSES = 64;
struct StreamFree {
  hipStream_t s;
  std::vector<void*> lists[ARENA_CLASSES];
};
struct LiveTable {
  // keys: block addresses (never null), values: class log2; linear probing
  // with backward-shift deletion, so no tombstones
  std::vector<void*> key;
  std::vector<unsigned char> cls;
  size_t n = 0;
  static size_t hash(void* p, size_t mask) {
    unsigned long long h = (unsigned long long) (size_t) p;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    return (size_t) h & mask;
  }
  void grow() {
    std::vector<void*> ok = std::move(key);
    std::vector<unsigned char> oc = std::move(cls);
    size_t cap = ok.empty() ? 1024 : ok.size() * 2;
    key.assign(cap, nullptr);
    cls.assign(cap, 0);
    n = 0;
    for (size_t i = 0; i < ok.size(); ++i)
      if (ok[i]) put(ok[i], oc[i]);
  }
  void put(void* p, unsigned char c) {
    if (2 * (n + 1) > key.size()) grow();
    size_t mask = key.size() - 1, i = hash(p, mask);
    while (key[i]) i = (i + 1) & mask;
    key[i] = p;
    cls[i] = c;
    ++n;
  }
  // the class of p, removed from the table; -1 if p is not live
  int take(void* p) {
    if (key.empty()) return -1;
    size_t mask = key.size() - 1, i = hash(p, mask);
    while (key[i] && key[i] != p) i = (i + 1) & mask;
    if (!key[i]) return -1;
    int c = cls[i];
    key[i] = nullptr;
    --n;
    // backward shift: move later entries of the cluster into the hole
    size_t hole = i;
    for (size_t j = (i + 1) & mask; key[j]; j = (j + 1) & mask) {
      size_t home = hash(key[j], mask);
      bool movable = hole <= j ? (home <= hole || home > j) : (home <= hole && home > j);
      if (movable) {
        key[hole] = key[j];
        cls[hole] = cls[j];
        key[j] = nullptr;
        hole = j;
      }
    }
    return c;
  }
};
struct Arena {
  std::mutex mu;
  std::vector<std::unique_ptr<StreamFree>> free_by_stream;
  LiveTable live;
  size_t in_use = 0, high = 0, cached = 0;
  StreamFree& of(hipStream_t s) {
    for (auto& f : free_by_stream)
      if (f->s == s) return *f;
    free_by_stream.push_back(std::make_unique<StreamFree>());
    free_by_stream.back()->s = s;
    return *free_by_stream.back();
  }
};
Arena& arena() {
  static Arena* a = new Arena;  // never destroyed: blocks may outlive static teardown
  return *a;
}
unsigned size_class_log2(size_t bytes) {
  unsigned k = 8;
  while (((size_t) 1 << k) < bytes) ++k;
  return k;
}
}  // namespace

void* dev_alloc(size_t bytes) {
  rt();
  hipStream_t s = stream();
  Arena& a = arena();
  unsigned const k = size_class_log2(bytes ? bytes : 1);
  size_t const c = (size_t) 1 << k;
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(a.mu);
    std::vector<void*>& fl = a.of(s).lists[k];
    if (!fl.empty()) {
      p = fl.back();
      fl.pop_back();
      a.cached -= c;
    }
  }
  if (!p) {
    auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      // give this stream's cached blocks back and retry once
      (void) hipGetLastError();
      dev_trim();
      OHX_CHECK(hipMalloc(&p, c));
    }
    // host-side count and wall time of the arena's misses
    if (prof_on())
      prof_host("host_malloc",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
  std::lock_guard<std::mutex> lock(a.mu);
  a.live.put(p, (unsigned char) k);
  a.in_use += c;
  if (a.in_use > a.high) a.high = a.in_use;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  Arena& a = arena();
  hipStream_t s = stream();
  std::lock_guard<std::mutex> lock(a.mu);
  int const k = a.live.take(p);
  if (k < 0) fail("dev_free of an unknown pointer", __FILE__, __LINE__);
  size_t const c = (size_t) 1 << k;
  a.in_use -= c;
  a.of(s).lists[k].push_back(p);
  a.cached += c;
}

void trim_stream(hipStream_t s) {
  Arena& a = arena();
  std::lock_guard<std::mutex> lock(a.mu);
  StreamFree& f = a.of(s);
  for (unsigned k = 0; k < ARENA_CLASSES; ++k) {
    for (void* p : f.lists[k]) {
      (void) hipFree(p);
      a.cached -= (size_t) 1 << k;
    }
    f.lists[k].clear();
  }
}

void dev_trim() {
  OHX_CHECK(hipStreamSynchronize(stream()));
  trim_stream(stream());
}

size_t dev_bytes_in_use() {
  std::lock_guard<std::mutex> lock(arena().mu);
  return arena().in_use;
}
size_t dev_bytes_high_water() {
  std::lock_guard<std::mutex> lock(arena().mu);
  return arena().high;
}
void dev_reset_high_water() {
  std::lock_guard<std::mutex> lock(arena().mu);
  arena().high = arena().in_use;
}

// ---------------------------------------------------------------- timing
namespace {
struct Prof {
  bool on = false;
  std::string only;  // when set, only the kernel of this name is timed
  std::vector<hipEvent_t> pool;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    double units;
  };
  std::vector<Pending> pending;
  std::unordered_map<std::string, hipEvent_t> open;
  struct Total {
    double ms = 0, units = 0;
    u64 launches = 0;
  };
  std::map<std::string, Total> totals;
};
Prof& prof() {
  static thread_local Prof p;  // each thread times the kernels it launches
  return p;
}
hipEvent_t take_event() {
  Prof& p = prof();
  if (!p.pool.empty()) {
    hipEvent_t e = p.pool.back();
    p.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  OHX_CHECK(hipEventCreate(&e));
  return e;
}
}  // namespace

bool prof_on() { return prof().on; }

bool prof_on_for(char const* name) {
  Prof& p = prof();
  return p.on && (p.only.empty() || p.only == name);
}

void prof_only(char const* name) { prof().only = name ? name : ""; }

void prof_host(char const* name, double ms) {
  auto& t = prof().totals[name];
  t.ms += ms;
  t.launches += 1;
}

ProfEvents prof_take() {
  ProfEvents e;
  e.a = take_event();
  e.b = take_event();
  return e;
}

void prof_record(char const* name, double bytes, ProfEvents e) {
  prof().pending.push_back({name, e.a, e.b, bytes});
}

void prof_enable(bool on) {
  Prof& p = prof();
  p.on = on;
}

void prof_collect() {
  Prof& p = prof();
  if (p.pending.empty()) return;
  sync();
  for (auto& q : p.pending) {
    float ms = 0;
    OHX_CHECK(hipEventElapsedTime(&ms, q.a, q.b));
    auto& t = p.totals[q.name];
    t.ms += ms;
    t.units += q.units;
    t.launches += 1;
    p.pool.push_back(q.a);
    p.pool.push_back(q.b);
  }
  p.pending.clear();
}

void prof_reset() {
  prof_collect();
  prof().totals.clear();
}

bool prof_get(char const* name, double* ms, u64* launches, double* units) {
  prof_collect();
  auto& t = prof().totals;
  auto it = t.find(name);
  if (it == t.end()) {
    *ms = 0;
    *launches = 0;
    if (units) *units = 0;
    return false;
  }
  *ms = it->second.ms;
  *launches = it->second.launches;
  if (units) *units = it->second.units;
  return true;
}

std::vector<std::string> prof_names() {
  prof_collect();
  std::vector<std::string> v;
  for (auto& kv : prof().totals) v.push_back(kv.first);
  return v;
}

}  // namespace ohx

namespace ohx {

static std::vector<Tuner*>& tuners() {
  static thread_local std::vector<Tuner*> v;
  return v;
}

// pinned variants are set between runs (not while threads adapt meshes)
static std::map<std::string, u32>& forced_variants() {
  static std::map<std::string, u32> m;
  return m;
}

void tuning_force(char const* name, int variant) {
  if (variant < 0)
    forced_variants().erase(name);
  else
    forced_variants()[name] = (u32) variant;
}

// Decisions measured earlier: tuner name -> size class -> variant.  A tuner
// takes the table's entry for a size class it has not timed yet instead of
// timing every variant there (each timing runs the whole kernel, which at
// large sizes costs as much as several passes save).  The table ships next
// to the library (tuning_gfx950.txt in the package directory, written from
// tuning reports of MI355X runs); OHX_TUNING_FILE names another file, or no
// file when empty; ohx_tuning_load adds entries.  Load before adapting.
static std::map<std::string, std::map<u32, u32>>& tuning_table() {
  static std::map<std::string, std::map<u32, u32>> t;
  return t;
}

// lines "name:2^class=variant" (the report's format; "?" entries skipped)
u32 tuning_load_text(std::string const& text) {
  u32 n = 0;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t end = text.find('\n', pos);
    if (end == std::string::npos) end = text.size();
    std::string line = text.substr(pos, end - pos);
    pos = end + 1;
    size_t colon = line.rfind(":2^"), eq = line.rfind('=');
    if (colon == std::string::npos || eq == std::string::npos || eq < colon) continue;
    char* stop = nullptr;
    unsigned long cls = std::strtoul(line.c_str() + colon + 3, &stop, 10);
    if (stop != line.c_str() + eq || cls > 32) continue;
    char const* vs = line.c_str() + eq + 1;
    if (*vs < '0' || *vs > '9') continue;
    unsigned long v = std::strtoul(vs, &stop, 10);
    tuning_table()[line.substr(0, colon)][(u32) cls] = (u32) v;
    ++n;
  }
  return n;
}

u32 tuning_load_file(char const* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) fail("tuning table: cannot open the file", __FILE__, __LINE__);
  std::string text;
  char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, k);
  std::fclose(f);
  return tuning_load_text(text);
}

static void load_default_table() {
  static std::once_flag once;
  std::call_once(once, [] {
    std::string path;
    if (char const* env = std::getenv("OHX_TUNING_FILE")) {
      path = env;
    } else {
      // <package>/lib/libomega_h_amd.so -> <package>/tuning_gfx950.txt
      Dl_info info;
      if (dladdr((void*) &load_default_table, &info) && info.dli_fname) {
        path = info.dli_fname;
        for (int up = 0; up < 2; ++up) {
          size_t slash = path.rfind('/');
          path = slash == std::string::npos ? std::string(".") : path.substr(0, slash);
        }
        path += "/tuning_gfx950.txt";
      }
    }
    if (path.empty()) return;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return;  // no table: every size class is timed online
    std::fclose(f);
    tuning_load_file(path.c_str());
  });
}

// timings per variant before a size class is decided (OHX_TUNING_TRIALS,
// default 1; more makes the choice robust to run-to-run noise on small
// kernels, for writing a tuning table)
static u32 tuning_trials() {
  static u32 const n = [] {
    char const* e = std::getenv("OHX_TUNING_TRIALS");
    long v = e ? std::strtol(e, nullptr, 10) : 1;
    return (u32) (v < 1 ? 1 : v > 64 ? 64 : v);
  }();
  return n;
}

static u32 size_class(u32 n) {
  u32 c = 0;
  while ((1u << c) < n && c < 32) ++c;
  return c;
}

void Tuner::harvest() {
  for (size_t k = 0; k < pending.size();) {
    Trial& tr = pending[k];
    if (hipEventQuery(tr.b) != hipSuccess) {
      ++k;
      continue;
    }
    float ms = 0;
    OHX_CHECK(hipEventElapsedTime(&ms, tr.a, tr.b));
    Cls& c = classes[tr.cls];
    double per = ms / (tr.items > 0 ? tr.items : 1);
    if (c.best[tr.v] < 0 || per < c.best[tr.v]) c.best[tr.v] = per;
    ++c.timed[tr.v];
    pool.push_back(tr.a);
    pool.push_back(tr.b);
    pending[k] = pending.back();
    pending.pop_back();
  }
  for (Cls& c : classes) {
    if (c.chosen >= 0 || c.best.empty()) continue;
    bool all = true;
    for (u32 t : c.timed) all = all && t >= tuning_trials();
    if (!all) continue;
    u32 arg = 0;
    for (u32 v = 1; v < c.best.size(); ++v)
      if (c.best[v] < c.best[arg]) arg = v;
    c.chosen = (int) arg;
  }
}

bool Tuner::forced() const { return forced_variants().count(name) != 0; }

u32 Tuner::pick(u32 n, hipEvent_t* a, hipEvent_t* b) {
  if (std::find(tuners().begin(), tuners().end(), this) == tuners().end()) tuners().push_back(this);
  auto f = forced_variants().find(name);
  if (f != forced_variants().end()) {
    for (u32 v : variants)
      if (v == f->second) return v;
  }
  u32 cls = size_class(n);
  Cls& c = classes[cls];
  if (c.best.empty()) {
    c.best.assign(variants.size(), -1.0);
    c.timed.assign(variants.size(), 0);
    load_default_table();
    auto t = tuning_table().find(name);
    if (t != tuning_table().end()) {
      auto e = t->second.find(cls);
      if (e != t->second.end())
        for (u32 k = 0; k < variants.size(); ++k)
          if (variants[k] == e->second) c.chosen = (int) k;
    }
  }
  if (c.chosen < 0 && !pending.empty()) harvest();
  if (c.chosen >= 0) return variants[c.chosen];
  // the next variant short of its timings (round robin while results are
  // due: a launch in flight counts once its events are read)
  u32 v = c.next % variants.size();
  for (u32 k = 0; k < variants.size(); ++k) {
    u32 cand = (c.next + k) % variants.size();
    if (c.timed[cand] < tuning_trials()) {
      v = cand;
      break;
    }
  }
  c.next = v + 1;
  for (int k = 0; k < 2; ++k) {
    hipEvent_t e;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else {
      OHX_CHECK(hipEventCreate(&e));
    }
    (k ? *b : *a) = e;
  }
  pending.push_back(Trial{cls, v, (double) n, *a, *b});
  return variants[v];
}

int Tuner::chosen(u32 n) const {
  int c = classes[size_class(n)].chosen;
  return c < 0 ? -1 : (int) variants[c];
}

std::string tuning_report() {
  std::string r;
  for (Tuner* t : tuners()) {
    t->harvest();
    for (u32 c = 0; c < t->classes.size(); ++c) {
      Tuner::Cls const& k = t->classes[c];
      if (k.best.empty()) continue;
      r += t->name;
      r += ":2^" + std::to_string(c) + "=";
      r += k.chosen < 0 ? std::string("?") : std::to_string(t->variants[k.chosen]);
      r += "\n";
    }
  }
  return r;
}

}  // namespace ohx
