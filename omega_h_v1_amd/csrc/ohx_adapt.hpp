// Adapt passes on a device-resident mesh.
#pragma once

#include "ohx_mesh.hpp"

namespace ohx {

Buf<u32> find_indset(Mesh& m, u32 ent_dim, u32 const* filter, double const* goodness);

u32 refine_by_size(Mesh& m, double qual_floor);
void uniformly_refine(Mesh& m);
u32 coarsen_by_size(Mesh& m, double qual_floor, double size_floor);
u32 coarsen_slivers(Mesh& m, double qual_floor, u32 nlayers);
u32 swap_slivers(Mesh& m, double good_qual, u32 nlayers);

struct AdaptStats {
  double npass = 0;   // passes that changed the mesh
  double nelems = 0;  // sum of their input element counts
  double tpass = 0;   // seconds inside those passes
  double ttotal = 0;  // seconds of the whole call
};
// returns 1 if any pass ran, 0 if none, -1 where the reference aborts
// (more than max_ops operations)
int adapt(Mesh& m, double size_floor, double good_qual, u32 nlayers, u32 max_ops, AdaptStats* st);
u32 warp_to_limit(Mesh& m, double qual_floor);

// swap tables (generated on first use)
u32 swap_nint_edges(u32 ring_size);
void swap_int_edges_host(u32 ring_size, u32 code, u32* out);

extern int g_verbose;
extern u32 g_indset_max_rounds;
extern bool g_mixed;
void mixed_stats(double* max_err, unsigned long long* screened, unsigned long long* reeval);
void mixed_stats_reset();

}  // namespace ohx
