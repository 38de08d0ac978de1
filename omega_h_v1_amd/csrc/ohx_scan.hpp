// Single-pass exclusive scan over a computed input, for any translation unit.
//
// exscan_fn(n, f, out, host_total): out[0] = 0, out[i+1] = f(0) + ... + f(i),
// where f is a __device__ functor; optionally the total goes to a host slot.
// Small inputs (one tile) use one block; larger ones a decoupled look-back
// scan (tile tickets, 64-bit (epoch, flag, value) tile words, bounded spin).
#pragma once

#include "ohx_common.hpp"

namespace ohx {

constexpr u32 SCAN_T = 256;
constexpr u32 SCAN_I = 16;
constexpr u32 SCAN_TILE = SCAN_T * SCAN_I;  // 4096 items per block

// tile words and ticket shared by every scan of the process
struct ScanTiles {
  u64* tiles;
  u64* ticket;
  u32 epoch;
};
ScanTiles scan_tiles_for(u32 ntiles);

namespace scan_detail {

__device__ inline u32 wave_incl_scan(u32 x) {
  u32 lane = threadIdx.x & 63u;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    u32 y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

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

__device__ inline u64 pack(u32 epoch, u32 inclusive, u32 v) {
  return ((u64) ((epoch << 1) | inclusive) << 32) | v;
}

__device__ inline u32 pad(u32 i) { return i + (i >> 5); }

template <class F>
__global__ __launch_bounds__(SCAN_T) void scan_fn_kernel(u32 n, F f, u32 ntiles, u32* __restrict__ out,
                                                         u64* __restrict__ tiles, u64* __restrict__ ticket,
                                                         u32 epoch, u32* __restrict__ host_total,
                                                         u32* __restrict__ err) {
  __shared__ u32 tile[SCAN_TILE + SCAN_TILE / 32];
  __shared__ u32 lds_w[SCAN_T / 64];
  __shared__ u32 s_tile, s_prefix;
  if (threadIdx.x == 0) {
    u32 t = 0;
    if (ntiles > 1) {
      // per-epoch ticket: the first block of a scan restarts the count
      u64 old = __hip_atomic_load(ticket, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (true) {
        u64 nxt;
        if ((u32) (old >> 32) != epoch) {
          nxt = ((u64) epoch << 32) | 1u;
          t = 0;
        } else {
          nxt = old + 1;
          t = (u32) old;
        }
        if (__hip_atomic_compare_exchange_strong(ticket, &old, nxt, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
          break;
      }
    }
    s_tile = t;
  }
  __syncthreads();
  u32 t = s_tile;
  if (t >= ntiles) {  // impossible with a consistent ticket; never spin on it
    if (threadIdx.x == 0) *err = ERR_SCAN_SPIN;
    return;
  }
  u32 base = t * SCAN_TILE;
#pragma unroll
  for (u32 k = 0; k < SCAN_I; ++k) {
    u32 li = k * SCAN_T + threadIdx.x;
    u32 idx = base + li;
    tile[pad(li)] = (idx < n) ? f(idx) : 0u;
  }
  __syncthreads();
  u32 mine[SCAN_I];
  u32 s = 0;
#pragma unroll
  for (u32 j = 0; j < SCAN_I; ++j) {
    mine[j] = tile[pad(threadIdx.x * SCAN_I + j)];
    s += mine[j];
  }
  u32 agg;
  u32 run = block_excl_scan(s, lds_w, &agg);
  if (threadIdx.x == 0) {
    u32 prefix = 0;
    if (ntiles > 1) {
      if (t == 0) {
        __hip_atomic_store(&tiles[0], pack(epoch, 1, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(&tiles[t], pack(epoch, 0, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        u32 j = t - 1;
        u32 spins = 0;
        while (true) {
          u64 v = __hip_atomic_load(&tiles[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          u32 hi = (u32) (v >> 32);
          if ((hi >> 1) != epoch) {
            if (++spins > (1u << 22)) {
              *err = ERR_SCAN_SPIN;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          prefix += (u32) v;
          if (hi & 1u) break;
          --j;
        }
        __hip_atomic_store(&tiles[t], pack(epoch, 1, prefix + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    s_prefix = prefix;
    if (t == ntiles - 1 && host_total) *host_total = prefix + agg;
  }
  __syncthreads();
  run += s_prefix;
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
  if (t == 0 && threadIdx.x == 0) out[0] = 0;
}

}  // namespace scan_detail

template <class F>
void exscan_fn(u32 n, F f, u32* out, u32* host_total_dev = nullptr) {
  u32 nt = n ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
  ScanTiles st = scan_tiles_for(nt);
  hipLaunchKernelGGL(scan_detail::scan_fn_kernel<F>, dim3(nt), dim3(SCAN_T), 0, stream(), n, f, nt, out,
                     st.tiles, st.ticket, st.epoch, host_total_dev, dev_error_word());
  check_launch("scan_fn_kernel");
}

// as exscan_fn and returns the total (one synchronisation)
template <class F>
u32 exscan_fn_total(u32 n, F f, u32* out) {
  exscan_fn(n, f, out, dev_slot<u32>(SLOT_SCAN));
  sync();
  return *host_slot<u32>(SLOT_SCAN);
}

}  // namespace ohx
