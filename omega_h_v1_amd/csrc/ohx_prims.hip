// Device-wide scans and reductions for gfx950 (64-lane wavefronts).
#include <cfloat>

#include "ohx_prims.hpp"

namespace ohx {

namespace {
constexpr u32 SCAN_T = 256;
constexpr u32 SCAN_I = 16;
constexpr u32 SCAN_TILE = SCAN_T * SCAN_I;  // 4096 items per block

__device__ inline u32 wave_incl_scan(u32 x) {
  u32 lane = threadIdx.x & 63u;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// exclusive scan of one value per thread over the 256-thread block
__device__ inline u32 block_excl_scan(u32 x, u32* lds_w, u32* total) {
  u32 incl = wave_incl_scan(x);
  u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 63u) lds_w[wave] = incl;
  __syncthreads();
  u32 off = 0, tot = 0;
#pragma unroll
  for (u32 w = 0; w < SCAN_T / 64; ++w) {
    u32 v = lds_w[w];
    if (w < wave) off += v;
    tot += v;
  }
  *total = tot;
  __syncthreads();
  return off + incl - x;
}

__global__ __launch_bounds__(SCAN_T) void scan_tile_sums(u32 const* a, u32 n, u32* sums) {
  __shared__ u32 lds_w[SCAN_T / 64];
  u32 base = blockIdx.x * SCAN_TILE;
  u32 s = 0;
#pragma unroll
  for (u32 k = 0; k < SCAN_I; ++k) {
    u32 idx = base + k * SCAN_T + threadIdx.x;
    if (idx < n) s += a[idx];
  }
  u32 tot;
  (void) block_excl_scan(s, lds_w, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// scans the tile of block b, adds offs[b] (if given), writes out[i+1]
__global__ __launch_bounds__(SCAN_T) void scan_tile_apply(u32 const* a, u32 n, u32 const* offs,
                                                          u32* out) {
  __shared__ u32 tile[SCAN_TILE + SCAN_TILE / 32];
  __shared__ u32 lds_w[SCAN_T / 64];
  u32 base = blockIdx.x * SCAN_TILE;
  auto pad = [](u32 i) { return i + (i >> 5); };
#pragma unroll
  for (u32 k = 0; k < SCAN_I; ++k) {
    u32 li = k * SCAN_T + threadIdx.x;
    u32 idx = base + li;
    tile[pad(li)] = (idx < n) ? a[idx] : 0u;
  }
  __syncthreads();
  u32 mine[SCAN_I];
  u32 s = 0;
#pragma unroll
  for (u32 j = 0; j < SCAN_I; ++j) {
    mine[j] = tile[pad(threadIdx.x * SCAN_I + j)];
    s += mine[j];
  }
  u32 tot;
  u32 run = block_excl_scan(s, lds_w, &tot) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
  for (u32 j = 0; j < SCAN_I; ++j) {
    run += mine[j];
    tile[pad(threadIdx.x * SCAN_I + j)] = run;
  }
  __syncthreads();
#pragma unroll
  for (u32 k = 0; k < SCAN_I; ++k) {
    u32 li = k * SCAN_T + threadIdx.x;
    u32 idx = base + li;
    if (idx < n) out[idx + 1] = tile[pad(li)];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

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
  Buf<T> part(nb + 1);
  hipLaunchKernelGGL((reduce_partial<T, Op>), dim3(nb), dim3(256), 0, stream(), a, begin, n, init,
                     part.get());
  hipLaunchKernelGGL((reduce_final<T, Op>), dim3(1), dim3(256), 0, stream(), part.get(), nb, init,
                     part.get() + nb);
  check_launch("reduce");
  return read_at(part.get(), nb);
}
}  // namespace

void exscan(u32 const* a, u32 n, u32* out) {
  if (n == 0) {
    fill(out, 1, 0);
    return;
  }
  u32 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    hipLaunchKernelGGL(scan_tile_apply, dim3(1), dim3(SCAN_T), 0, stream(), a, n, nullptr, out);
    check_launch("scan_tile_apply");
    return;
  }
  Buf<u32> sums(nb);
  Buf<u32> offs(nb + 1);
  hipLaunchKernelGGL(scan_tile_sums, dim3(nb), dim3(SCAN_T), 0, stream(), a, n, sums.get());
  check_launch("scan_tile_sums");
  exscan(sums.get(), nb, offs.get());
  hipLaunchKernelGGL(scan_tile_apply, dim3(nb), dim3(SCAN_T), 0, stream(), a, n, offs.get(), out);
  check_launch("scan_tile_apply");
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
  Buf<u32> neg(n);
  u32* g = neg.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { g[i] = (offs[i + 1] == offs[i]) ? 1u : 0u; });
  return exscan(g, n);
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
  // allocated once for the life of the process; never returned to the arena
  static u32* s = static_cast<u32*>(dev_alloc(64 * sizeof(u32)));
  return s;
}

}  // namespace ohx
