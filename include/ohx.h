/* Kernel-level C ABI and mesh seam of the MI355X adapt library.
 *
 * Two layers sit below include/omega_h.h:
 *
 * 1. ohx_* kernel entry points: plain device pointers, sizes and a HIP
 *    stream (passed as void*, NULL = the library stream).  Each replaces one
 *    data-parallel loop of the reference and cites it.  They enqueue work on
 *    the stream; entry points with a host return value synchronise.  On a
 *    bad argument or a device error they return OHX_ERROR (or a NaN for
 *    double results) and ohx_last_error() says why.
 *
 * 2. ohx_mesh_* seam: the reference's internal pass functions (refine.hpp,
 *    coarsen.hpp, swap.hpp, adapt.hpp, warp_to_limit.hpp, derive_model.hpp)
 *    over a device-resident mesh, with host arrays in and out.  Tests and
 *    the benchmark drive the product through this layer, in the same shape
 *    as the oracle's orc_* and the reference harness's ref_* functions.
 *
 * There is no CPU fallback: every call needs a visible HIP device.
 */
#ifndef OHX_H
#define OHX_H

#include "omega_h.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OHX_ERROR 0xffffffffu

/* last error message of this thread, or NULL */
char const* ohx_last_error(void) OSH_PUBLIC;
void ohx_clear_error(void) OSH_PUBLIC;

/* ------------------------------------------------- kernel entry points */

/* elem_quality_kern / element_qualities (quality.cpp:13-45): FP64 quality
 * of every element (tets: Shewchuk V/A_rms^{3/2}; triangles: signed z-area
 * form), bit-identical to the reference's x86-64 build. */
unsigned ohx_element_qualities_f64(unsigned dim, unsigned nelems,
    unsigned const* verts_of_elems, double const* coords, double* out,
    void* stream) OSH_PUBLIC;

/* min_element_quality (quality.cpp:47-57) with the serial doubles_min
 * (doubles.cpp:111-118): the minimum over elements 1..nelems-1. */
double ohx_min_quality_f64(unsigned dim, unsigned nelems,
    unsigned const* verts_of_elems, double const* coords, void* stream) OSH_PUBLIC;

/* measure_edge (size.cpp:57-88): |x1-x0| / ((h0+h1)/2) per edge */
unsigned ohx_measure_edges_f64(unsigned nedges, unsigned const* verts_of_edges,
    double const* coords, double const* size, double* out, void* stream) OSH_PUBLIC;

/* uints_exscan (ints.cpp:25-36): out[0]=0, out[i+1]=a[0]+...+a[i] (n+1 values) */
unsigned ohx_exscan_u32(unsigned const* a, unsigned n, unsigned* out, void* stream) OSH_PUBLIC;

/* No counterpart in the reference: ohx_exscan_u32 for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It allocates nothing: its
 * look-back tile words live in a fixed device array of the library (256 KiB
 * per GPU, held until the process ends), so n is limited to 2^27 =
 * 134217728 items; a larger n returns OHX_ERROR with ohx_last_error() and
 * launches nothing.  Calls on one GPU share that array (it keeps the last
 * call's contents), so calls that overlap on one GPU (two host threads, or
 * two streams with no ordering between them) race.  stream: the caller's
 * stream (NULL: the null stream), synchronised before the call returns. */
unsigned ohx_exscan_u32_noalloc(unsigned const* a, unsigned n, unsigned* out, void* stream) OSH_PUBLIC;

/* up_from_down (up_from_down.cpp:24-47) over invert_map
 * (invert_map.cpp:118-136): offsets (nlows+1), highs and directions
 * (nhighs*deg each) sorted by (high, direction) per low. */
unsigned ohx_up_from_down(unsigned high_dim, unsigned low_dim, unsigned nhighs,
    unsigned nlows, unsigned const* lows_of_highs, unsigned* offsets,
    unsigned* highs_of_lows, unsigned* directions, void* stream) OSH_PUBLIC;

/* find_indset (indset.cpp:95-117) on a CSR graph: state[i] = 1 for members
 * of the independent set.  Returns the number of Jacobi rounds run (in
 * batches), or OHX_ERROR if the set did not converge within 100 rounds. */
unsigned ohx_find_indset(unsigned n, unsigned const* offsets, unsigned const* adj,
    unsigned const* filter, double const* goodness, unsigned* state,
    void* stream) OSH_PUBLIC;

/* FP64 self-test of the device arithmetic the parity claims rest on:
 * sqrt_out[i] = sqrt(a[i]), div_out[i] = a[i] / b[i], computed with the
 * library's compile flags, to be compared bitwise with the host. */
unsigned ohx_fp64_selftest(unsigned n, double const* a, double const* b,
    double* sqrt_out, double* div_out, void* stream) OSH_PUBLIC;

/* Anisotropic extension (no counterpart in the reference, whose
 * size_from_hessian.cpp:12-40 builds only a scalar size): FP64 cyclic Jacobi
 * eigendecomposition of symmetric 3x3 tensors (6 components m00 m01 m02 m11
 * m12 m22 each): lam (3 per tensor), V (3x3 row-major per tensor, columns
 * are eigenvectors) ... */
unsigned ohx_eigen3_f64(unsigned n, double const* h6, double* lam, double* V9,
    void* stream) OSH_PUBLIC;
/* ... and the metric M = V diag(clamp(|lam|/eps, 1/hmax^2, 1/hmin^2)) V^T */
unsigned ohx_metric_from_hessian_f64(unsigned n, double const* h6, double eps,
    double hmin, double hmax, double* m6, void* stream) OSH_PUBLIC;

/* ----------------------------------------------------------- mesh seam */
typedef struct ohx_mesh ohx_mesh;

/* a reduced (vertices + elements) serial mesh from host arrays */
ohx_mesh* ohx_mesh_new(unsigned dim, unsigned nverts, double const* coords,
    unsigned nelems, unsigned const* conn) OSH_PUBLIC;
/* Kuhn box: n^dim cubes (squares) split as new_box_mesh (mesh.cpp:69-83,
 * tables.cpp:15-47), generated on the device */
ohx_mesh* ohx_mesh_new_box(unsigned dim, unsigned n) OSH_PUBLIC;
ohx_mesh* ohx_mesh_clone(ohx_mesh* m) OSH_PUBLIC;
void ohx_mesh_free(ohx_mesh* m) OSH_PUBLIC;
/* mesh_derive_model(m, PI/4) + mesh_set_rep(m, MESH_FULL)
 * (derive_model.cpp:444-448, test_warp.cpp:108-110) */
void ohx_mesh_derive_model_full(ohx_mesh* m) OSH_PUBLIC;
void ohx_mesh_set_full(ohx_mesh* m) OSH_PUBLIC;

unsigned ohx_mesh_count(ohx_mesh* m, unsigned dim) OSH_PUBLIC;
void ohx_mesh_get_down(ohx_mesh* m, unsigned h, unsigned l, unsigned* out) OSH_PUBLIC;
unsigned ohx_mesh_up_nuses(ohx_mesh* m, unsigned l, unsigned h) OSH_PUBLIC;
void ohx_mesh_get_up(ohx_mesh* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj, unsigned* dirs) OSH_PUBLIC;
unsigned ohx_mesh_star_nadj(ohx_mesh* m, unsigned l, unsigned h) OSH_PUBLIC;
void ohx_mesh_get_star(ohx_mesh* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj) OSH_PUBLIC;
void ohx_mesh_get_dual(ohx_mesh* m, unsigned* out) OSH_PUBLIC;
/* returns ncomps, or -1 if there is no such tag of that kind */
int ohx_mesh_get_tag(ohx_mesh* m, unsigned dim, char const* name, int want_f64,
    void* out) OSH_PUBLIC;
unsigned ohx_mesh_ntags(ohx_mesh* m, unsigned dim) OSH_PUBLIC;
char const* ohx_mesh_tag_name(ohx_mesh* m, unsigned dim, unsigned i) OSH_PUBLIC;
void ohx_mesh_add_field(ohx_mesh* m, unsigned dim, char const* name,
    unsigned ncomps, double const* data) OSH_PUBLIC;
/* osh_new_field (omega_h.cpp:613-622) with its data: an FP64 field whose
 * enum osh_transfer type (1 pointwise, 2 conserve) makes an element field
 * follow refine, coarsen and swap */
void ohx_mesh_add_field_tt(ohx_mesh* m, unsigned dim, char const* name,
    unsigned ncomps, double const* data, int transfer) OSH_PUBLIC;
void ohx_mesh_free_field(ohx_mesh* m, unsigned dim, char const* name) OSH_PUBLIC;
/* the size-field producers of the reference's warp loop (test_warp.cpp:
 * 70-83): mesh_element_gradients (element_gradients.cpp:66-89) adds the
 * element field "grad_<name>" (3 components per component of the vertex
 * field name); mesh_recover_by_volume (recover_by_volume.cpp:52-66) adds
 * the vertex field name, the size-weighted mean of the element field name;
 * mesh_size_from_hessian (size_from_hessian.cpp:44-72) adds the vertex
 * field "adapt_size" = max_h - sum_j weights[j] |H_j|_F, at least min_h
 * (weights: one per 9 components, or NULL for all 1); and
 * mesh_identity_size_field (size.cpp:90-135), the longest edge at each
 * vertex, as osh_identity_size */
void ohx_mesh_element_gradients(ohx_mesh* m, char const* name) OSH_PUBLIC;
void ohx_mesh_recover_by_volume(ohx_mesh* m, char const* name) OSH_PUBLIC;
void ohx_mesh_size_from_hessian(ohx_mesh* m, char const* hess_name,
    double const* weights, double min_h, double max_h) OSH_PUBLIC;
void ohx_mesh_identity_size(ohx_mesh* m, char const* name) OSH_PUBLIC;
/* uniform field on every vertex, set on the device */
void ohx_mesh_set_uniform_field(ohx_mesh* m, char const* name, double value) OSH_PUBLIC;

/* a vertex "metric" field (6 components) from a vertex Hessian field; a
 * mesh with a "metric" field adapts in that metric instead of "adapt_size" */
void ohx_mesh_metric_from_hessian(ohx_mesh* m, char const* hessian, double eps,
    double hmin, double hmax, char const* out_name) OSH_PUBLIC;

double ohx_mesh_min_quality(ohx_mesh* m) OSH_PUBLIC;
void ohx_mesh_qualities(ohx_mesh* m, double* out) OSH_PUBLIC;
void ohx_mesh_measure_edges(ohx_mesh* m, double* out) OSH_PUBLIC;
void ohx_mesh_find_indset(ohx_mesh* m, unsigned ent_dim,
    unsigned const* candidates, double const* goodness, unsigned* out) OSH_PUBLIC;

/* refine.cpp:17-35, coarsen.cpp:23-70, swap.cpp:138-154 */
unsigned ohx_mesh_refine_by_size(ohx_mesh* m, double qual_floor) OSH_PUBLIC;
void ohx_mesh_uniformly_refine(ohx_mesh* m) OSH_PUBLIC;
unsigned ohx_mesh_coarsen_by_size(ohx_mesh* m, double qual_floor,
    double size_floor) OSH_PUBLIC;
unsigned ohx_mesh_coarsen_slivers(ohx_mesh* m, double qual_floor,
    unsigned nlayers) OSH_PUBLIC;
unsigned ohx_mesh_swap_slivers(ohx_mesh* m, double good_qual,
    unsigned nlayers) OSH_PUBLIC;
/* mesh_adapt (adapt.cpp:82-94); -1 where the reference aborts after
 * max_ops operations */
int ohx_mesh_adapt(ohx_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops) OSH_PUBLIC;
/* as ohx_mesh_adapt; stats[0] passes that changed the mesh, [1] the sum of
 * their input element counts, [2] seconds inside them, [3] seconds total */
int ohx_mesh_adapt_timed(ohx_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops, double* stats) OSH_PUBLIC;
/* mesh_warp_to_limit (warp_to_limit.cpp:61-81) */
unsigned ohx_mesh_warp_to_limit(ohx_mesh* m, double qual_floor) OSH_PUBLIC;

/* Batches (no counterpart in the reference, which adapts one mesh per call):
 * n full, classified meshes of one dimension, carrying the same fields,
 * become one batch mesh (the inputs are left as they are).  Adapting the
 * batch runs one launch sequence for all of them and leaves every component
 * exactly as mesh_adapt would leave it alone; results[c] is that call's
 * return value (1, 0, or -1 at the operation limit), stats as in
 * ohx_mesh_adapt_timed (input elements count the components a pass changed).
 * Returns the number of components that reached the operation limit. */
ohx_mesh* ohx_mesh_batch_new(unsigned n, ohx_mesh* const* meshes) OSH_PUBLIC;
unsigned ohx_mesh_batch_size(ohx_mesh* m) OSH_PUBLIC;
/* component c as a mesh of its own (entities in batch order, own class ids) */
ohx_mesh* ohx_mesh_batch_extract(ohx_mesh* m, unsigned c) OSH_PUBLIC;
/* A pass costs the whole batch, so during ohx_mesh_batch_adapt the finished
 * components are set aside once they hold more than half of the batch's
 * elements and the batch has at least this many (default 1e6); the result
 * is the same batch either way. */
void ohx_set_batch_retire_min(double elements) OSH_PUBLIC;
/* When the batch pass that would serve the most elements covers less than
 * this fraction of the batch (default 0.25), the live components finish
 * their sequences one mesh at a time; 0 never hands over, 1 always does.
 * The result is the same batch either way. */
void ohx_set_batch_handoff(double fraction) OSH_PUBLIC;
int ohx_mesh_batch_adapt(ohx_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops, int* results, double* stats) OSH_PUBLIC;

/* mesh_adapt on n independent meshes at once (no counterpart in the
 * reference): worker threads (at least nthreads; they persist) each adapt
 * meshes on a HIP stream of their own.  results[i] is ohx_mesh_adapt's
 * value for meshes[i]; stats[0..2] sum the meshes' passes, input elements
 * and pass seconds, stats[3] is the wall time of the call.  Returns the
 * number of meshes that reached the operation limit, or -2 with
 * ohx_last_error() if any mesh failed. */
int ohx_adapt_many(unsigned n, ohx_mesh* const* meshes, double size_floor,
    double good_qual, unsigned nlayers, unsigned max_ops, unsigned nthreads,
    int* results, double* stats) OSH_PUBLIC;

/* read_mesh_vtk / write_mesh_vtk (vtk_io.cpp:787-830): the .vtu piece (and
 * with a .pvtu name the .pvtu index too) in the reference's format; write
 * returns 0, or -1 with ohx_last_error() */
ohx_mesh* ohx_mesh_read_vtk(char const* path) OSH_PUBLIC;
int ohx_mesh_write_vtk(ohx_mesh* m, char const* path) OSH_PUBLIC;
/* 1: write arrays zlib-compressed (as a reference build with zlib does);
 * 0 (default): uncompressed, as a reference build without zlib, readable by
 * either.  Reading handles both. */
void ohx_set_vtk_compression(int on) OSH_PUBLIC;

/* mesh_diff (mesh_diff.cpp:141-178): 0 if the meshes have the same dimension,
 * entity counts, connectivity and tags (integers exactly, FP64 to relative
 * tolerance tol above the absolute floor), else 1 after printing what
 * differs, as the reference does; with allow_superset the second mesh may
 * hold extra tags and dimensions.  OHX_ERROR on failure. */
unsigned ohx_mesh_diff(ohx_mesh* a, ohx_mesh* b, double tol, double floor,
    unsigned allow_superset) OSH_PUBLIC;

/* the osh_t handle of a seam mesh (same object) */
osh_t ohx_mesh_osh(ohx_mesh* m) OSH_PUBLIC;

/* ------------------------------------------------------ instrumentation */
void ohx_set_verbose(int on) OSH_PUBLIC;
/* round limit of the independent set: 100 as the reference (indset.cpp:33,
 * the default; 0 restores it); a higher limit is an extension for meshes
 * whose equal-goodness candidates form long chains (each round settles at
 * least the lowest-numbered candidate of a tie chain) */
void ohx_set_indset_max_rounds(unsigned n) OSH_PUBLIC;
/* the most rounds any independent set needed since the last reset, among
 * the sets that took more than one batch of 8 rounds (0 if none did);
 * reset != 0 clears it after reading */
unsigned ohx_indset_rounds_max(int reset) OSH_PUBLIC;
/* 0: the FP64 path (default, the reference's arithmetic); 1: FP32-mixed --
 * qualities and metric lengths are screened in FP32 and every value within
 * a margin of a decision (a threshold or a competing value) is re-evaluated
 * in FP64, so every decision equals the FP64 path's (extension) */
void ohx_set_precision(int mixed) OSH_PUBLIC;
/* FP32-mixed statistics since the last reset: the largest |q32 - q64| seen
 * where both were computed, decisions settled in FP32, FP64 re-evaluations */
void ohx_mixed_stats(double* max_err, unsigned long long* screened,
    unsigned long long* reeval) OSH_PUBLIC;
void ohx_mixed_stats_reset(void) OSH_PUBLIC;
/* the online autotuner's choices so far, "kernel:2^class=variant" per line
 * (variant: the kernel's own code, e.g. lanes per item, items per thread,
 * 0/1 for a plain or combined form); returns the full length */
unsigned ohx_tuning_report(char* buf, unsigned cap) OSH_PUBLIC;
/* pin the tuner of that name (as in the report) to one of its variants;
 * variant < 0 lets it tune again.  Every variant gives identical results. */
void ohx_tuning_force(char const* name, int variant) OSH_PUBLIC;
/* add earlier decisions, lines in the report's format, to the table the
 * tuners consult before timing a size class (the library loads
 * tuning_gfx950.txt from its package directory, or the file OHX_TUNING_FILE
 * names, at its first tuned launch); returns the entries read, OHX_ERROR if
 * the file cannot be read.  Call before adapting. */
unsigned ohx_tuning_load(char const* path) OSH_PUBLIC;
/* per-kernel timing: while enabled, each timed launch carries a start and a
 * stop event in its own dispatch (hipExtLaunchKernel), so the interval is
 * the kernel's execution alone */
void ohx_profile_enable(int on) OSH_PUBLIC;
/* restrict the timing to the kernel of this name (NULL: every kernel) */
void ohx_profile_only(char const* name) OSH_PUBLIC;
void ohx_profile_reset(void) OSH_PUBLIC;
/* total milliseconds, launches and algorithmic HBM bytes (each input read
 * once, each output written once) of a timed kernel; returns 0 if unseen */
int ohx_profile_get(char const* name, double* ms, unsigned long long* launches,
    double* units) OSH_PUBLIC;
/* newline-separated names of the timed kernels; returns the full length */
unsigned ohx_profile_names(char* buf, unsigned cap) OSH_PUBLIC;
unsigned long long ohx_device_bytes_high_water(void) OSH_PUBLIC;
void ohx_device_reset_high_water(void) OSH_PUBLIC;
void ohx_synchronize(void) OSH_PUBLIC;
/* interior-edge table of the swap triangulations (edge_swap_edges.cpp) */
unsigned ohx_swap_nint_edges(unsigned ring_size) OSH_PUBLIC;
void ohx_swap_int_edges(unsigned ring_size, unsigned code, unsigned* out) OSH_PUBLIC;

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif
