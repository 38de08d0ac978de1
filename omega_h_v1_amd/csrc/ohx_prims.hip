// Device-wide scans and reductions for gfx950 (64-lane wavefronts).
#include <cfloat>

#include <atomic>

#include "ohx_prims.hpp"
#include "ohx_scan.hpp"

namespace ohx {

namespace {
template <class T>
struct MinOp {
  __device__ static T apply(T a, T b) { return b < a ? b : a; }
};
template <class T>
struct MaxOp {
  __device__ static T apply(T a, T b) { return b > a ? b : a; }
};
template <class T>
struct SumOp {
  __device__ static T apply(T a, T b) { return a + b; }
};

template <class T, class Op>
__device__ inline T block_reduce(T v, T* lds) {
#pragma unroll
  for (u32 d = 32; d >= 1; d >>= 1) v = Op::apply(v, __shfl_down(v, d, 64));
  u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T r = lds[0];
    for (u32 w = 1; w < blockDim.x / 64; ++w) r = Op::apply(r, lds[w]);
    lds[0] = r;
  }
  __syncthreads();
  T r = lds[0];
  __syncthreads();
  return r;
}

constexpr u32 RED_BLOCKS = 512;

template <class T, class Op>
__global__ __launch_bounds__(256) void reduce_partial(T const* a, u32 begin, u32 n, T init,
                                                      T* part) {
  __shared__ T lds[4];
  T v = init;
  for (u32 i = begin + blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    v = Op::apply(v, a[i]);
  v = block_reduce<T, Op>(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

template <class T, class Op>
__global__ __launch_bounds__(256) void reduce_final(T const* part, u32 np, T init, T* out) {
  __shared__ T lds[4];
  T v = init;
  for (u32 i = threadIdx.x; i < np; i += 256) v = Op::apply(v, part[i]);
  v = block_reduce<T, Op>(v, lds);
  if (threadIdx.x == 0) *out = v;
}

template <class T, class Op>
T reduce(T const* a, u32 n, u32 begin, T init) {
  if (n <= begin) return init;
  u32 nb = nblocks(n - begin, 256);
  if (nb > RED_BLOCKS) nb = RED_BLOCKS;
  Buf<T> part(nb);
  hipLaunchKernelGGL((reduce_partial<T, Op>), dim3(nb), dim3(256), 0, stream(), a, begin, n, init,
                     part.get());
  hipLaunchKernelGGL((reduce_final<T, Op>), dim3(1), dim3(256), 0, stream(), part.get(), nb, init,
                     dev_slot<T>(SLOT_REDUCE));
  check_launch("reduce");
  sync();
  return *host_slot<T>(SLOT_REDUCE);
}
}  // namespace

namespace {
// per thread (the thread's stream): tile words, never cleared between scans
// (each scan tags its words with a new epoch)
struct ScanState {
  u64* tiles = nullptr;
  u32* ticket = nullptr;
  u32 cap = 0;
  u32 epoch = 0;
};
std::atomic<u32> g_scan_threads{0};
ScanState& scan_state() {
  static thread_local ScanState st;
  if (!st.ticket) {
    // the thread's ticket word, zero between scans (the last block of a
    // ticketed scan clears it); a thread that scans stays counted, since
    // driving threads live until the process ends
    st.ticket = static_cast<u32*>(dev_alloc(sizeof(u32)));
    OHX_CHECK(hipMemsetAsync(st.ticket, 0, sizeof(u32), stream()));
    g_scan_threads.fetch_add(1);
  }
  return st;
}
}  // namespace

bool scan_concurrent() { return g_scan_threads.load(std::memory_order_relaxed) > 1; }

ScanTiles scan_tiles_for(u32 nt) {
  ScanState& st = scan_state();
  if (nt > st.cap) {
    // a larger tile array; the old one is released once the stream drains
    if (st.tiles) {
      sync();
      dev_free(st.tiles);
    }
    st.cap = nt < 1024 ? 1024 : nt;
    st.tiles = static_cast<u64*>(dev_alloc(sizeof(u64) * st.cap));
    OHX_CHECK(hipMemsetAsync(st.tiles, 0xff, sizeof(u64) * st.cap, stream()));
  }
  // epochs stay below 2^30 - 1, so fresh all-ones words never match one
  st.epoch = st.epoch + 1 >= 0x3fffffffu ? 1u : st.epoch + 1;
  return ScanTiles{st.tiles, st.ticket, st.epoch};
}

void exscan(u32 const* a, u32 n, u32* out) {
  exscan_fn(n, OHX_LAMBDA(u32 i) { return a[i]; }, out);
}

u32 exscan_total(u32 const* a, u32 n, u32* out) {
  return exscan_fn_total(n, OHX_LAMBDA(u32 i) { return a[i]; }, out);
}

Buf<u32> exscan(u32 const* a, u32 n) {
  Buf<u32> o(n + 1);
  exscan(a, n, o.get());
  return o;
}

u32 max_u32(u32 const* a, u32 n) { return reduce<u32, MaxOp<u32>>(a, n, 0, 0u); }
u32 sum_u32(u32 const* a, u32 n) { return reduce<u32, SumOp<u32>>(a, n, 0, 0u); }
double min_f64(double const* a, u32 n, u32 begin) {
  return reduce<double, MinOp<double>>(a, n, begin, DBL_MAX);
}

namespace {
__global__ __launch_bounds__(256) void min_final_both(double const* part, u32 np, double* dev_out,
                                                      double* host_out) {
  __shared__ double lds[4];
  double v = DBL_MAX;
  for (u32 i = threadIdx.x; i < np; i += 256) v = MinOp<double>::apply(v, part[i]);
  v = block_reduce<double, MinOp<double>>(v, lds);
  if (threadIdx.x == 0) {
    *dev_out = v;
    *host_out = v;
  }
}
}  // namespace

void min_f64_async(double const* a, u32 n, u32 begin, double* dev_out) {
  u32 nb = n > begin ? nblocks(n - begin, 256) : 0;
  if (nb > RED_BLOCKS) nb = RED_BLOCKS;
  Buf<double> part(nb ? nb : 1);
  if (nb)
    hipLaunchKernelGGL((reduce_partial<double, MinOp<double>>), dim3(nb), dim3(256), 0, stream(), a, begin, n,
                       DBL_MAX, part.get());
  hipLaunchKernelGGL(min_final_both, dim3(1), dim3(256), 0, stream(), part.get(), nb, dev_out,
                     dev_slot<double>(SLOT_MIN_ASYNC));
  check_launch("min_f64_async");
}

double min_f64_async_result() { return *host_slot<double>(SLOT_MIN_ASYNC); }
double max_f64(double const* a, u32 n, u32 begin) {
  return reduce<double, MaxOp<double>>(a, n, begin, -DBL_MAX);
}

void fill(u32* a, u32 n, u32 v) {
  parallel_for(n, OHX_LAMBDA(u32 i) { a[i] = v; });
}

Buf<u32> filled(u32 n, u32 v) {
  Buf<u32> b(n);
  fill(b.get(), n, v);
  return b;
}

Buf<double> filled_f64(u32 n, double v) {
  Buf<double> b(n);
  double* p = b.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { p[i] = v; });
  return b;
}

Buf<u32> iota(u32 n) {
  Buf<u32> b(n);
  u32* p = b.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { p[i] = i; });
  return b;
}

Buf<u32> copy(u32 const* a, u32 n) {
  Buf<u32> b(n);
  d2d(b.get(), a, n);
  return b;
}

Buf<double> copy_f64(double const* a, u32 n) {
  Buf<double> b(n);
  d2d(b.get(), a, n);
  return b;
}

Buf<u32> negate_offsets(u32 const* offs, u32 n) {
  // offsets of 0/1 flags: the items before i without a product number
  // i - offs[i], so no scan is needed
  Buf<u32> o(n + 1);
  u32* op = o.get();
  parallel_for(n + 1, OHX_LAMBDA(u32 i) { op[i] = i - offs[i]; });
  return o;
}

Buf<u32> scale(u32 const* a, u32 n, u32 s) {
  Buf<u32> b(n);
  u32* p = b.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { p[i] = s * a[i]; });
  return b;
}

void expand_into_u32(u32* out, u32 const* a, u32 const* offs, u32 n, u32 width) {
  parallel_for(n, OHX_LAMBDA(u32 i) {
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      for (u32 k = 0; k < width; ++k) out[j * width + k] = a[i * width + k];
  });
}

void expand_into_f64(double* out, double const* a, u32 const* offs, u32 n, u32 width) {
  parallel_for(n, OHX_LAMBDA(u32 i) {
    for (u32 j = offs[i]; j < offs[i + 1]; ++j)
      for (u32 k = 0; k < width; ++k) out[j * width + k] = a[i * width + k];
  });
}

void compact_u32(u32* out, u32 const* a, u32 const* offs, u32 n, u32 width) {
  u32 total = n * width;
  parallel_for(total, OHX_LAMBDA(u32 t) {
    u32 i = t / width, k = t - i * width;
    u32 o = offs[i];
    if (offs[i + 1] != o) out[o * width + k] = a[t];
  });
}

u32* scratch_words() {
  // one per thread, for the life of the thread
  static thread_local u32* w = nullptr;
  if (!w) {
    w = static_cast<u32*>(dev_alloc(64 * sizeof(u32)));
    fill(w, 64, 0);
  }
  return w;
}

}  // namespace ohx
