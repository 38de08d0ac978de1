// Device-wide primitives: scans, reductions, fills and gathers over HBM
// arrays.  All run on the library stream; functions that return a host
// value synchronize once.
#pragma once

#include "ohx_common.hpp"

namespace ohx {

// out[0] = 0, out[i+1] = a[0] + ... + a[i]; out has n+1 entries
void exscan(u32 const* a, u32 n, u32* out);
Buf<u32> exscan(u32 const* a, u32 n);
// as exscan, and returns the total out[n] to the host (one synchronisation)
u32 exscan_total(u32 const* a, u32 n, u32* out);
// reductions returning a host scalar
u32 max_u32(u32 const* a, u32 n);
u32 sum_u32(u32 const* a, u32 n);
// min / max over a[1..n-1]: the serial loop build starts at index 1
// (doubles.cpp:100-118); begin selects the first index
double min_f64(double const* a, u32 n, u32 begin);
// the minimum of a[begin, n) (DBL_MAX if empty) into dev_out and into a
// mapped host slot, with no wait: min_f64_async_result() holds it after the
// stream's next synchronisation
void min_f64_async(double const* a, u32 n, u32 begin, double* dev_out);
double min_f64_async_result();
double max_f64(double const* a, u32 n, u32 begin);

Buf<u32> filled(u32 n, u32 v);
Buf<double> filled_f64(u32 n, double v);
void fill(u32* a, u32 n, u32 v);
Buf<u32> iota(u32 n);
Buf<u32> copy(u32 const* a, u32 n);
Buf<double> copy_f64(double const* a, u32 n);
// uints_negate_offsets (ints.cpp:224-232) for offsets whose steps are 0 or 1
// (the exclusive scan of 0/1 flags): out[i] = i - offs[i]
Buf<u32> negate_offsets(u32 const* offs, u32 n);
// uints_scale
Buf<u32> scale(u32 const* a, u32 n, u32 s);
// expand_array (arrays.cpp:187-213) for 4- and 8-byte elements
void expand_into_u32(u32* out, u32 const* a, u32 const* offs, u32 n, u32 width);
void expand_into_f64(double* out, double const* a, u32 const* offs, u32 n, u32 width);
// compaction: out[offs[i]..] = a[i*w..] for entities with offs[i] != offs[i+1],
// where each source entity maps to at most one slot (the common case of
// expand); uses a thread per (entity, component)
void compact_u32(u32* out, u32 const* a, u32 const* offs, u32 n, u32 width);

// small device scratch of 64 words for flags/counters (one per runtime)
u32* scratch_words();

}  // namespace ohx
