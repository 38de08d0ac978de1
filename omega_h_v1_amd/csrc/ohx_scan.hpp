// Single-pass exclusive scan over a computed input, for any translation unit.
//
// exscan_fn(n, f, out, host_total): out[0] = 0, out[i+1] = f(0) + ... + f(i),
// where f is a __device__ functor; optionally the total goes to a host slot.
// Small inputs (one tile) use one block; larger ones a decoupled look-back
// scan (tiles in block order, 64-bit (epoch, flag, value) tile words, a
// wave-wide look-back, bounded spin).  Tiles hold 256 x I items; I (4 or
// 16) and the look-back window (64 or 256 tiles per step) are chosen per
// call site and size class by the online tuner: small scans want more,
// smaller tiles to reach every CU and a window that spans them, large ones
// fewer tiles and a narrow window.
#pragma once

#include "ohx_common.hpp"

namespace ohx {

constexpr u32 SCAN_T = 256;

// tile words shared by every scan of a host thread (its stream), and the
// thread's tile ticket word (zero between scans)
struct ScanTiles {
  u64* tiles;
  u32* ticket;
  u32 epoch;
};
ScanTiles scan_tiles_for(u32 ntiles);
// true once more than one host thread drives the library: scans then take
// their tile from a ticket (see scan_fn_kernel)
bool scan_concurrent();

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

template <u32 SCAN_I, u32 LB, bool TICKET, class F>
__global__ __launch_bounds__(SCAN_T) void scan_fn_kernel(u32 n, F f, u32 ntiles, u32* __restrict__ out,
                                                         u64* __restrict__ tiles, u32* __restrict__ ticket,
                                                         u32 epoch, u32* __restrict__ host_total,
                                                         u32* __restrict__ err) {
  constexpr u32 SCAN_TILE = SCAN_T * SCAN_I;
  __shared__ u32 tile[SCAN_TILE + SCAN_TILE / 32];
  __shared__ u32 lds_w[SCAN_T / 64];
  __shared__ u32 s_prefix;
  __shared__ u32 s_tile;
  // Which tile a block scans.  With one host thread (one stream of scans)
  // the tile is the block index: workgroups are dispatched in index order
  // within each XCD, so the lowest-numbered resident block always finds
  // every earlier tile finished or on an XCD with free slots, and the
  // look-back cannot wait on a block that is never scheduled.  That
  // argument does not cover scans of other streams holding the free slots,
  // so with several host threads (TICKET) a block takes its tile from a
  // fetch-add ticket: tiles are then claimed only by resident blocks, in
  // order, and the lowest unfinished tile always belongs to a running block
  // whose predecessors are all done, whatever else occupies the GPU.  (The
  // ticket costs one same-address atomic per block, which retire one per
  // ~12 ns on this part, so the single-stream path does without it.)  The
  // block that draws the last ticket clears the word for the next scan of
  // the stream.  The spin stays bounded regardless.
  u32 t = blockIdx.x;
  if (TICKET) {
    if (threadIdx.x == 0) {
      u32 k = atomicAdd(ticket, 1u);
      if (k == ntiles - 1) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_tile = k;
    }
    __syncthreads();
    t = s_tile;
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
  if (threadIdx.x < 64) {
    // wave 0 publishes this tile and looks back over up to 64 x LB
    // predecessor tiles per step (lane l reads tiles j - LB l .. j - LB l -
    // LB + 1): the closest inclusive prefix ends the walk once every tile
    // before it in the window is published.  A wide window (LB = 4) reaches
    // an inclusive prefix in one step on scans of up to 256 tiles; on large
    // scans its extra polling loads compete with the tiles' own traffic in
    // L2, so the tuner picks LB per size class.
    u32 lane = threadIdx.x;
    u32 prefix = 0;
    if (ntiles > 1) {
      if (lane == 0)
        __hip_atomic_store(&tiles[t], pack(epoch, t == 0 ? 1u : 0u, agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (t > 0) {
        int j = (int) t - 1;
        u32 spins = 0;
        while (true) {
          u64 v[LB];
#pragma unroll
          for (u32 q = 0; q < LB; ++q) {
            int idx = j - (int) (LB * lane + q);
            v[q] = idx >= 0 ? __hip_atomic_load(&tiles[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : pack(epoch, 1, 0);
          }
          // per lane: the first inclusive entry (LB if none) and whether an
          // entry before it (or any entry, if none) is still unpublished
          u32 qi = LB;
          bool waiting = false;
#pragma unroll
          for (u32 q = 0; q < LB; ++q) {
            u32 hi = (u32) (v[q] >> 32);
            bool ready = (hi >> 1) == epoch;
            if (qi == LB) {
              if (!ready) waiting = true;
              else if (hi & 1u) qi = q;
            }
          }
          u64 pmask = __ballot(qi < LB);
          u32 first = pmask ? (u32) (__ffsll((long long) pmask) - 1) : 64u;
          // lanes up to the first with an inclusive entry must be complete
          bool need = lane <= first;
          if (__ballot(need && waiting)) {
            if (++spins > (1u << 22)) {
              if (lane == 0) *err = ERR_SCAN_SPIN;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          u32 val = 0;
          if (need) {
            u32 upto = lane < first ? LB : qi + 1;
#pragma unroll
            for (u32 q = 0; q < LB; ++q)
              if (q < upto) val += (u32) v[q];
          }
#pragma unroll
          for (u32 d = 32; d >= 1; d >>= 1) val += __shfl_xor(val, d, 64);
          prefix += val;
          if (first < 64) break;
          j -= 64 * (int) LB;
        }
        if (lane == 0)
          __hip_atomic_store(&tiles[t], pack(epoch, 1, prefix + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (lane == 0) {
      s_prefix = prefix;
      if (t == ntiles - 1 && host_total) *host_total = prefix + agg;
    }
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

template <u32 I, u32 LB, bool TICKET, class F>
void exscan_launch_t(u32 n, F f, u32* out, u32* host_total_dev, hipEvent_t ea, hipEvent_t eb) {
  constexpr u32 tile = SCAN_T * I;
  u32 nt = n ? (n + tile - 1) / tile : 1;
  ScanTiles st = scan_tiles_for(nt);
  if (ea) {
    hipExtLaunchKernelGGL(scan_detail::scan_fn_kernel<I, LB, TICKET, F>, dim3(nt), dim3(SCAN_T), 0, stream(), ea, eb,
                          0, n, f, nt, out, st.tiles, st.ticket, st.epoch, host_total_dev, dev_error_word());
    check_launch("scan");
    return;
  }
  // compulsory bytes: one 4-byte input per item (a lower bound for functors
  // that read more) and the n + 1 offsets out
  double prof_bytes = (double) n * 4 + (double) (n + 1) * 4;
  launch_timed("scan", prof_bytes, scan_detail::scan_fn_kernel<I, LB, TICKET, F>, dim3(nt), dim3(SCAN_T), n, f, nt,
               out, st.tiles, st.ticket, st.epoch, host_total_dev, dev_error_word());
}

template <u32 I, u32 LB, class F>
void exscan_launch(u32 n, F f, u32* out, u32* host_total_dev, hipEvent_t ea, hipEvent_t eb) {
  if (scan_concurrent())
    exscan_launch_t<I, LB, true>(n, f, out, host_total_dev, ea, eb);
  else
    exscan_launch_t<I, LB, false>(n, f, out, host_total_dev, ea, eb);
}

template <class F>
void exscan_fn(u32 n, F f, u32* out, u32* host_total_dev = nullptr) {
  // one tuner per call site (functor type) and host thread; variants: 4 or
  // 16 items per thread with a 64-tile look-back window, or (44) 4 items
  // with a 256-tile window
  static thread_local Tuner tuner("scan", {4, 16, 44});
  tuned(tuner, n, [&](u32 v, hipEvent_t ea, hipEvent_t eb) {
    if (v == 4)
      exscan_launch<4, 1>(n, f, out, host_total_dev, ea, eb);
    else if (v == 44)
      exscan_launch<4, 4>(n, f, out, host_total_dev, ea, eb);
    else
      exscan_launch<16, 1>(n, f, out, host_total_dev, ea, eb);
  });
}

// as exscan_fn and returns the total (one synchronisation)
template <class F>
u32 exscan_fn_total(u32 n, F f, u32* out) {
  exscan_fn(n, f, out, dev_slot<u32>(SLOT_SCAN));
  sync();
  return *host_slot<u32>(SLOT_SCAN);
}

// exscan_fn for a functor whose items each take several dependent memory
// round trips (a loop over an entity's edges, hash probes).  Inside the scan
// kernel a thread evaluates its 4 or 16 items one after another, so those
// round trips add up per thread; here the items are first evaluated one per
// thread into a buffer, where the whole grid overlaps them, and the buffer
// is then scanned: one more launch and 8 bytes per item of traffic.  Below
// one scan tile the functor is scanned as it is.
template <class F>
void exscan_fn_heavy(u32 n, F f, u32* out) {
  if (n < (1u << 12)) {
    exscan_fn(n, f, out);
    return;
  }
  Buf<u32> tmp(n);
  u32* t = tmp.get();
  parallel_for(n, OHX_LAMBDA(u32 i) { t[i] = f(i); });
  exscan_fn(n, OHX_LAMBDA(u32 i) { return t[i]; }, out);
}

}  // namespace ohx
