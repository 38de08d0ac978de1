// The exclusive scan of the kernel ABI for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): ohx_exscan_u32_noalloc.
//
// ohx_exscan_u32 takes its look-back tile words from the library's arena.
// This variant keeps them in a fixed device array of this file, scratch,
// of NA_WORDS 64-bit words (256 KiB + 16 bytes), which every GPU the library
// runs on holds until the process ends.  In host code the array's name is
// only a stand-in for the symbol: its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that
// address is used, for the kernel arguments and the clearing kernel.
//
// Layout: NA_MAX_TILES tile words (epoch-tagged, never cleared between
// calls; each call uses a new epoch), then the tile ticket word (zero
// between calls: the block that draws the last ticket clears it), then the
// error word of the look-back's spin bound.  The scan is the ticketed form
// of the library's single-pass scan, so it stays live whatever else runs
// on the GPU.  All calls on one GPU share the array and keep the last
// call's contents: calls that overlap on one GPU race.
#include <mutex>

#include "../../include/ohx.h"
#include "ohx_scan.hpp"

namespace ohx {
void api_set_error(char const* what);  // ohx_api.hip: ohx_last_error()
void noalloc_setup_current_device();
}  // namespace ohx

using namespace ohx;

namespace {
constexpr u32 NA_ITEMS = 16;                          // items per thread
constexpr u32 NA_TILE = SCAN_T * NA_ITEMS;            // 4096 items per tile
constexpr u32 NA_MAX_TILES = 32768;                   // 256 KiB of tile words
constexpr u64 NA_MAX_ITEMS = (u64) NA_MAX_TILES * NA_TILE;  // 2^27 items
constexpr u32 NA_TICKET = NA_MAX_TILES;               // word index of the ticket
constexpr u32 NA_ERR = NA_MAX_TILES + 1;              // word index of the error word
constexpr u32 NA_WORDS = NA_MAX_TILES + 2;
constexpr u32 NA_EPOCH_LIMIT = 0x3fffffffu;           // epochs fit 30 bits
constexpr int NA_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) u64 scratch[NA_WORDS];

namespace {
struct NaDevice {
  u64* ptr = nullptr;  // device address of scratch on this device
  u32 epoch = 0;       // last epoch used on this device
};
std::mutex g_na_mu;
NaDevice g_na[NA_MAX_DEVICES];

__global__ __launch_bounds__(BLOCK) void na_clear(u64* __restrict__ w, u32 n) {
  u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) w[i] = 0ull;
}

struct LoadU32 {
  u32 const* a;
  __device__ u32 operator()(u32 i) const { return a[i]; }
};

void na_check(hipError_t e, char const* what) {
  if (e != hipSuccess) fail(what, __FILE__, __LINE__);
}

// once per device, with that device current: the array's address, and the
// whole array cleared (zero tile words match no epoch, zero ticket)
NaDevice& na_device(hipStream_t s) {
  int dev = 0;
  na_check(hipGetDevice(&dev), "hipGetDevice");
  OHX_REQUIRE(dev >= 0 && dev < NA_MAX_DEVICES, "ohx_exscan_u32_noalloc: device index out of range");
  NaDevice& d = g_na[dev];
  if (!d.ptr) {
    u64* p = nullptr;
    na_check(hipGetSymbolAddress((void**) &p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    hipLaunchKernelGGL(na_clear, dim3(nblocks(NA_WORDS)), dim3(BLOCK), 0, s, p, NA_WORDS);
    na_check(hipGetLastError(), "na_clear launch");
    d.ptr = p;
    d.epoch = 0;
  }
  return d;
}
}  // namespace

namespace ohx {
void noalloc_setup_current_device() {
  std::lock_guard<std::mutex> lock(g_na_mu);
  na_device(nullptr);
  na_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}
}  // namespace ohx

extern "C" unsigned ohx_exscan_u32_noalloc(unsigned const* a, unsigned n, unsigned* out, void* stream) {
  try {
    if ((u64) n > NA_MAX_ITEMS)
      fail("ohx_exscan_u32_noalloc: n exceeds the fixed scratch (at most 2^27 = 134217728 items)", __FILE__,
           __LINE__);
    hipStream_t s = static_cast<hipStream_t>(stream);
    u64* p;
    u32 epoch;
    {
      std::lock_guard<std::mutex> lock(g_na_mu);
      NaDevice& d = na_device(s);
      if (d.epoch + 1 >= NA_EPOCH_LIMIT) {
        // epochs wrap: clear the tile words so no stale word can match
        hipLaunchKernelGGL(na_clear, dim3(nblocks(NA_MAX_TILES)), dim3(BLOCK), 0, s, d.ptr, NA_MAX_TILES);
        na_check(hipGetLastError(), "na_clear launch");
        d.epoch = 0;
      }
      d.epoch += 1;
      p = d.ptr;
      epoch = d.epoch;
    }
    u32 nt = n ? (n + NA_TILE - 1) / NA_TILE : 1;
    u32* ticket = reinterpret_cast<u32*>(p + NA_TICKET);
    u32* err = reinterpret_cast<u32*>(p + NA_ERR);
    hipLaunchKernelGGL((scan_detail::scan_fn_kernel<NA_ITEMS, 1, true, LoadU32>), dim3(nt), dim3(SCAN_T), 0, s, n,
                       LoadU32{a}, nt, out, p, ticket, epoch, static_cast<u32*>(nullptr), err);
    na_check(hipGetLastError(), "scan launch");
    na_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    u32 e = 0;
    na_check(hipMemcpy(&e, err, sizeof(u32), hipMemcpyDeviceToHost), "hipMemcpy (error word)");
    if (e) {
      hipLaunchKernelGGL(na_clear, dim3(1), dim3(BLOCK), 0, s, p + NA_ERR, 1u);
      na_check(hipGetLastError(), "na_clear launch");
      na_check(hipStreamSynchronize(s), "hipStreamSynchronize");
      fail("scan: look-back did not complete (spin limit)", __FILE__, __LINE__);
    }
    return 0u;
  } catch (std::exception const& ex) {
    api_set_error(ex.what());
    return OHX_ERROR;
  }
}
