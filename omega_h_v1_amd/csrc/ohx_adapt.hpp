// Adapt passes on a device-resident mesh.
#pragma once

#include <atomic>

#include "ohx_mesh.hpp"

namespace ohx {

Buf<u32> find_indset(Mesh& m, u32 ent_dim, u32 const* filter, double const* goodness);

u32 refine_by_size(Mesh& m, double qual_floor);
void uniformly_refine(Mesh& m);
// The size codes of the collapse-by-size loop, per edge of the current
// mesh: a surviving edge keeps its vertices, their coordinates and the
// metric, so its code (its length against the floor) stays what it was;
// only the product edges of a pass are measured again.
struct SizeCodeCache {
  u32 ne = 0;  // edges of the mesh it describes (0: empty)
  Buf<u32> codes;
  void clear() {
    ne = 0;
    codes.release();
  }
};
u32 coarsen_by_size(Mesh& m, double qual_floor, double size_floor, SizeCodeCache* cache = nullptr);
// eq: the mesh's element qualities when the caller has them (else computed)
u32 coarsen_slivers(Mesh& m, double qual_floor, u32 nlayers, double const* eq = nullptr);
// The swap choices of the edges evaluated by earlier swap passes of one
// shape loop, per edge of the current mesh: an edge whose tets are the
// same keeps its ring, so its swap quality, ring size, triangulation and
// acceptance stay what they were (info: bit 31 valid, bit 30 accepted,
// bits 16-23 ring size, bits 0-15 triangulation).  Any other pass that
// changes the mesh clears it.
struct SwapCache {
  u32 ne = 0;  // edges of the mesh it describes (0: empty)
  Buf<double> q;
  Buf<u32> info;
  // the new mesh's element qualities after a swap pass that changed it
  // (survivors' carried, products' evaluated), for the shape loop to take
  Buf<double> eq;
  void clear() {
    ne = 0;
    q.release();
    info.release();
    eq.release();
  }
};
u32 swap_slivers(Mesh& m, double good_qual, u32 nlayers, double const* eq = nullptr, double const* gate = nullptr,
                 SwapCache* cache = nullptr);

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

// ------------------------------------------------------------- batching
// A batch is one mesh holding many independent meshes (components): each
// component's entities keep their relative order through every pass, and
// its class ids are shifted into a range of its own per class dimension, so
// the component of an entity is found from its class tags.  A batch pass
// runs on the components the driver marks active, with each component's own
// quality floor, and reports which components it changed; the driver steps
// every component through mesh_adapt's sequence as if it were alone.
struct BatchCtx {
  u32 ncomp = 0;
  Buf<u32> off;      // 4 rows of ncomp + 1: first class id of each component per class dim
  Buf<u32> active;   // per component: the current pass may change it
  Buf<double> qf;    // per component: quality floor of the size passes
  Buf<u32> changed;  // per component: the current pass changed it
};
extern thread_local BatchCtx* g_batch;
// a batch retires its finished components once they hold more than half of
// its elements and it has at least this many elements
extern double g_batch_retire_min;
// a batch hands its live components to single-mesh runs when the best batch
// pass would change components holding less than this fraction of it
extern double g_batch_handoff;

__device__ inline u32 batch_comp(u32 const* off, u32 ncomp, u32 cd, u32 cid) {
  u32 const* row = off + (size_t) cd * (ncomp + 1);
  u32 lo = 0, hi = ncomp;  // row[lo] <= cid < row[hi] (row[ncomp] is all ones)
  while (hi - lo > 1) {
    u32 mid = (lo + hi) / 2;
    if (row[mid] <= cid)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
// zero flags[e] of entities of dimension d in inactive components
void batch_mask(Mesh& m, u32 d, u32* flags);
// the quality floor of each entity's component (dimension d)
Buf<double> batch_floors(Mesh& m, u32 d);
// record the components owning a set entity of dimension d
void batch_note_changed(Mesh& m, u32 d, u32 const* indset);

// the union of classified full meshes of one dimension, as one mesh
void batch_union(std::vector<Mesh*> const& parts, Mesh& out, BatchCtx& b);
// component c of a batch as a mesh of its own (entities in batch order)
void batch_extract(Mesh& u, BatchCtx& b, u32 c, Mesh& out);
// per-component element counts and min qualities (the serial skip of each
// component's first element included)
void batch_quality(Mesh& m, BatchCtx& b, std::vector<double>& minq, std::vector<double>& nelems);
// mesh_adapt on every component; results[c] as adapt()'s return value
void batch_adapt(Mesh& m, BatchCtx& b, double size_floor, double good_qual, u32 nlayers, u32 max_ops, int* results,
                 AdaptStats* st);

extern int g_verbose;
extern u32 g_indset_max_rounds;
extern std::atomic<u32> g_indset_rounds_seen;
extern bool g_mixed;
void mixed_stats(double* max_err, unsigned long long* screened, unsigned long long* reeval);
void mixed_stats_reset();

// solution transfer of the element fields tagged OSH_TRANSFER_CONSERVE /
// OSH_TRANSFER_POINTWISE, called by each operation once its output
// elements (and their class tags) exist
void refine_transfer(Mesh& m, Mesh& mo, u32 const* soff, u32 const* off, u32 ppd, u32 go, u32 nout);
void coarsen_transfer(Mesh& m, Mesh& mo, u32 const* vgoff, u32 ncol, u32 const* egoff, u32 const* esoff, u32 nsame,
                      u32 ngen);
void swap_transfer(Mesh& m, Mesh& mo, u32 const* egoff, u32 const* soff, u32 nsame, u32 ngen);

}  // namespace ohx
