// Common definitions of the MI355X (gfx950) adapt-path library.
//
// One process, one GPU, one HIP stream per library instance.  Every mesh
// array lives in HBM; the host keeps only counts and control flow.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ohx {

using u32 = unsigned;
using u64 = unsigned long long;
constexpr u32 INVALID = ~0u;
constexpr u32 MAX_UP = 128;
constexpr u32 MAX_EDGE_SWAP = 7;
constexpr double REF_PI = 3.14159265359;  // algebra.hpp:8

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void fail(char const* what, char const* file, int line);
#define OHX_CHECK(x)                                                       \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess)                                                  \
      ::ohx::fail(hipGetErrorString(e_), __FILE__, __LINE__);              \
  } while (0)
#define OHX_REQUIRE(c, msg)                                                \
  do {                                                                     \
    if (!(c)) ::ohx::fail(msg, __FILE__, __LINE__);                        \
  } while (0)

// --------------------------------------------------------------- runtime
// Per host thread: the library's stream and a small pinned staging area for
// the scalars the host reads back (scan totals, reduction results,
// convergence flags).  Threads driving different meshes therefore run on
// streams of their own and never share scratch state.
struct Runtime {
  hipStream_t stream = nullptr;
  void* pinned = nullptr;      // 4 KiB of mapped pinned host memory ...
  void* pinned_dev = nullptr;  // ... and its device address
  int device = 0;
};
Runtime& rt();
// a caller-provided stream for the duration of one kernel-ABI call
extern thread_local hipStream_t g_stream_override;
inline hipStream_t stream() { return g_stream_override ? g_stream_override : rt().stream; }
struct StreamScope {
  hipStream_t prev;
  bool own;
  // s == nullptr: the call runs on the library stream and completes before
  // returning (the caller has no handle to wait on); otherwise it runs on
  // the caller's stream, which is drained at the end because the call's
  // temporaries go back to the arena, which reuses blocks in library-stream
  // order
  explicit StreamScope(hipStream_t s) : prev(g_stream_override), own(s == nullptr) {
    if (s) g_stream_override = s;
  }
  ~StreamScope() {
    (void) hipStreamSynchronize(stream());
    g_stream_override = prev;
  }
};
void sync();

// Host-visible slots (8 bytes each) in mapped pinned memory: kernels write
// the few scalars the host needs (scan totals, reductions, error codes)
// straight into them, so a read costs one stream synchronisation and no
// copy.  Slot 0 is the device error word, checked by every sync().
enum Slot : unsigned { SLOT_ERROR = 0, SLOT_SCAN = 1, SLOT_REDUCE = 2, SLOT_MIN_ASYNC = 3, SLOT_GATHER = 8, SLOT_STAGE = 32 };
template <class T>
inline T* host_slot(unsigned k) { return reinterpret_cast<T*>(static_cast<char*>(rt().pinned) + 8 * k); }
template <class T>
inline T* dev_slot(unsigned k) { return reinterpret_cast<T*>(static_cast<char*>(rt().pinned_dev) + 8 * k); }
inline unsigned* dev_error_word() { return dev_slot<unsigned>(SLOT_ERROR); }
// device error codes (set by kernels, reported by the next sync())
enum DevError : unsigned {
  ERR_NONE = 0,
  ERR_REFLECT_NOT_FOUND = 1,
  ERR_DUAL_NONMANIFOLD = 2,
  ERR_HINGE_SIDES = 3,
  ERR_REFINE_QUALITY = 4,
  ERR_EXPOSED_SIDE = 5,
  ERR_NO_DESTINATION = 6,
  ERR_EDGE_RING = 7,
  ERR_SCAN_SPIN = 8,
};

// ------------------------------------------------------------------ arena
// Caching allocator over hipMalloc: power-of-two size classes, free lists
// reused in stream order (all work is on one stream, so a block freed by
// the host may be handed to the next launch without an event).
void* dev_alloc(size_t bytes);
void dev_free(void* p);
size_t dev_bytes_in_use();
size_t dev_bytes_high_water();
void dev_reset_high_water();
void dev_trim();

template <class T>
class Buf {
 public:
  Buf() = default;
  // always allocates (like malloc(0) in the reference), so an existing
  // but empty array stays distinguishable from an absent one
  explicit Buf(size_t n) : n_(n) { p_ = static_cast<T*>(dev_alloc(n ? n * sizeof(T) : sizeof(T))); }
  ~Buf() { release(); }
  Buf(Buf const&) = delete;
  Buf& operator=(Buf const&) = delete;
  Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void release() {
    if (p_) dev_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// -------------------------------------------------------- host <-> device
template <class T>
void to_host(T* dst, T const* src, size_t n) {
  if (!n) return;
  OHX_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream()));
  sync();
}
template <class T>
Buf<T> to_device(T const* src, size_t n) {
  Buf<T> b(n);
  if (n) {
    OHX_CHECK(hipMemcpyAsync(b.get(), src, n * sizeof(T), hipMemcpyHostToDevice, stream()));
    sync();
  }
  return b;
}
template <class T>
std::vector<T> host_copy(T const* src, size_t n) {
  std::vector<T> v(n);
  to_host(v.data(), src, n);
  return v;
}
// read one device scalar (one D2H round trip)
template <class T>
T read_at(T const* p, size_t i) {
  T* h = host_slot<T>(SLOT_STAGE);
  OHX_CHECK(hipMemcpyAsync(h, p + i, sizeof(T), hipMemcpyDeviceToHost, stream()));
  sync();
  return *h;
}
// read several device scalars with one launch and one synchronisation
void read_many(unsigned const* const* ptrs, unsigned n, unsigned* out);
template <class T>
void d2d(T* dst, T const* src, size_t n) {
  if (!n) return;
  OHX_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));
}

// ---------------------------------------------------------------- launch
constexpr u32 BLOCK = 256;

// The dispatcher hands consecutive workgroups to the 8 XCDs in turn, and
// each XCD has its own L2.  xcd_block() can permute the block index so that
// each XCD works on one contiguous eighth of the items (a bijection on
// [0, gridDim.x)).  Measured on the adapt kernels (1.6 M and 24.6 M tet boxes) that was
// 10-20% slower than the plain order -- neighbouring blocks on different XCDs
// share the Infinity Cache and spread over more HBM channels at once -- so
// the permutation is compiled out unless OHX_XCD_REMAP=1.
constexpr u32 NXCD = 8;
#ifndef OHX_XCD_REMAP
#define OHX_XCD_REMAP 0
#endif
__device__ inline u32 xcd_block() {
  if (!OHX_XCD_REMAP) return blockIdx.x;
  u32 b = blockIdx.x, n = gridDim.x;
  u32 x = b % NXCD, k = b / NXCD;
  u32 per = n / NXCD, rem = n % NXCD;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}
inline u32 nblocks(size_t n, u32 block = BLOCK) { return static_cast<u32>((n + block - 1) / block); }
void check_launch(char const* name);

template <class F>
__global__ __launch_bounds__(BLOCK) void for_kernel(u32 n, F f) {
  u32 i = xcd_block() * BLOCK + threadIdx.x;
  if (i < n) f(i);
}
// one thread per index; the body is a __device__ lambda capturing by value
template <class F>
void parallel_for(u32 n, F f) {
  if (!n) return;
  hipLaunchKernelGGL(for_kernel<F>, dim3(nblocks(n)), dim3(BLOCK), 0, stream(), n, f);
  check_launch("parallel_for");
}
#define OHX_LAMBDA [=] __device__

// *w = 1 once per wave if any active lane has pred: stores of many lanes to
// one word serialise in L2 and hold up the kernel's completion
__device__ inline void wave_flag(u32* w, bool pred) {
  unsigned long long m = __ballot(pred);
  unsigned long long act = __ballot(1);
  if (m && (u32) __lane_id() == (u32) (__ffsll((long long) act) - 1)) *w = 1u;
}
// *w = v once per wave if any active lane has pred (a tagged "any" word: the
// host compares it with the v of this launch, so it needs no clearing)
__device__ inline void wave_flag_value(u32* w, bool pred, u32 v) {
  unsigned long long m = __ballot(pred);
  unsigned long long act = __ballot(1);
  if (m && (u32) __lane_id() == (u32) (__ffsll((long long) act) - 1)) *w = v;
}
// *w = 1 for the lanes with pred, one store per distinct address in the wave
__device__ inline void wave_flag_at(u32* w, bool pred) {
  unsigned long long todo = __ballot(pred);
  while (todo) {
    int leader = __ffsll((long long) todo) - 1;
    unsigned long long addr = (unsigned long long) __shfl((long long) (size_t) w, leader);
    bool same = pred && (unsigned long long) (size_t) w == addr;
    if ((int) __lane_id() == leader) *w = 1u;
    todo &= ~__ballot(same);
  }
}
// atomicAdd(&counts[key], 1) for the lanes with pred, returning each lane's
// old value as if the adds ran one after another in some order: the lanes
// with the same key are combined, one atomic per distinct key in the wave,
// and a lane's value is the leader's returned base plus its rank among them.
// Every lane of the wave must call it (it ballots).
__device__ inline u32 wave_slot_add(u32* counts, u32 key, bool pred) {
  unsigned long long todo = __ballot(pred);
  u32 const lane = (u32) __lane_id();
  unsigned long long const below = (1ull << lane) - 1ull;
  u32 leader_of = lane, rank = 0, n = 0;
  while (todo) {
    int leader = __ffsll((long long) todo) - 1;
    u32 lk = (u32) __builtin_amdgcn_readlane((int) key, leader);
    bool same = pred && key == lk;
    unsigned long long mm = __ballot(same);
    if (same) {
      leader_of = (u32) leader;
      rank = (u32) __popcll(mm & below);
    }
    if (lane == (u32) leader) n = (u32) __popcll(mm);
    todo &= ~mm;
  }
  u32 base = 0;
  if (pred && leader_of == lane) base = atomicAdd(&counts[key], n);
  base = (u32) __shfl((int) base, (int) leader_of);
  return base + rank;
}

// ---------------------------------------------------------------- timing
// Per-kernel-name timers, enabled by ohx_profile_enable().  A timed launch
// carries its start and stop events in its own dispatch
// (hipExtLaunchKernel), so the interval is the kernel's execution alone, as
// a kernel trace measures it; bytes: the launch's algorithmic (compulsory)
// HBM bytes, summed per name.
struct ProfEvents {
  hipEvent_t a = nullptr, b = nullptr;
};
bool prof_on();
bool prof_on_for(char const* name);
void prof_only(char const* name);
// a host-side interval (e.g. time spent waiting in a synchronisation)
void prof_host(char const* name, double ms);
ProfEvents prof_take();
void prof_record(char const* name, double bytes, ProfEvents e);

template <class K, class... A>
void launch_timed(char const* name, double bytes, K kernel, dim3 grid, dim3 block, A... args) {
  if (prof_on_for(name)) {
    ProfEvents e = prof_take();
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream(), e.a, e.b, 0, args...);
    check_launch(name);
    prof_record(name, bytes, e);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, stream(), args...);
    check_launch(name);
  }
}

// ---------------------------------------------------------------- autotune
// Online choice among launch variants that compute identical results (e.g.
// lanes per item).  Per size class (log2 of the item count) every variant
// is tried on live launches with start/stop events in its dispatch; once
// each has been timed, later launches of that class use the fastest (time
// per item).  Timings are read only when their events have completed, so
// tuning adds no synchronisation.
struct Tuner {
  char const* name;
  std::vector<u32> variants;
  struct Trial {
    u32 cls, v;
    double items;
    hipEvent_t a, b;
  };
  struct Cls {
    std::vector<double> best;  // ms per item per variant, < 0 if untimed
    std::vector<u32> timed;    // timings harvested per variant
    int chosen = -1;
    u32 next = 0;
  };
  std::vector<Cls> classes = std::vector<Cls>(33);
  std::vector<Trial> pending;
  std::vector<hipEvent_t> pool;
  Tuner(char const* n, std::vector<u32> v) : name(n), variants(std::move(v)) {}
  // the variant for n items; events to put in the dispatch (null if none)
  u32 pick(u32 n, hipEvent_t* a, hipEvent_t* b);
  void harvest();
  int chosen(u32 n) const;
  // pinned to a variant by tuning_force
  bool forced() const;
};
// run launch(variant, start_event, stop_event) with the tuner's choice
template <class L>
void tuned(Tuner& t, u32 n, L launch) {
  hipEvent_t a = nullptr, b = nullptr;
  u32 v = t.pick(n, &a, &b);
  launch(v, a, b);
}
// tuning decisions so far, "name:class=variant" per line
std::string tuning_report();
// pin a tuner to one of its variants (variant < 0: tune again)
void tuning_force(char const* name, int variant);
// add decisions ("name:2^class=variant" lines) to the table the tuners
// consult before timing; returns the number of entries read
u32 tuning_load_text(std::string const& text);
u32 tuning_load_file(char const* path);

}  // namespace ohx
