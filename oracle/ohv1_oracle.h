/* CPU oracle for the Omega_h v1 adapt-pass hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch serial C restatement of
 * the reference's algorithm (FP64, isotropic "adapt_size" metric, serial
 * loop build semantics).  Only tests/, bench.py's cpu_baseline leg and
 * __graft_entry__.smoke() load it, and only as the checker.  The product
 * (omega_h_v1_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef OHV1_ORACLE_H
#define OHV1_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh orc_mesh;

/* Kuhn box generator: n^dim cubes (squares), each split like the
 * reference's single box (tables.cpp:29-47). Caller frees with orc_free_buf. */
void orc_box(unsigned dim, unsigned n, unsigned* nverts, double** coords,
    unsigned* nelems, unsigned** conn);
void orc_free_buf(void* p);

orc_mesh* orc_new_mesh(unsigned dim, unsigned nverts, double const* coords,
    unsigned nelems, unsigned const* conn);
void orc_free(orc_mesh* m);
void orc_derive_model_full(orc_mesh* m);
void orc_derive_model(orc_mesh* m, double crease_angle);
void orc_set_full(orc_mesh* m);

unsigned orc_dim(orc_mesh* m);
unsigned orc_count(orc_mesh* m, unsigned dim);
void orc_get_down(orc_mesh* m, unsigned h, unsigned l, unsigned* out);
unsigned orc_up_nuses(orc_mesh* m, unsigned l, unsigned h);
void orc_get_up(orc_mesh* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj, unsigned* dirs);
unsigned orc_star_nadj(orc_mesh* m, unsigned l, unsigned h);
void orc_get_star(orc_mesh* m, unsigned l, unsigned h, unsigned* offs,
    unsigned* adj);
void orc_get_dual(orc_mesh* m, unsigned* out);
int orc_get_tag(orc_mesh* m, unsigned dim, char const* name, int want_f64,
    void* out);
unsigned orc_ntags(orc_mesh* m, unsigned dim);
char const* orc_tag_name(orc_mesh* m, unsigned dim, unsigned i);
void orc_add_field(orc_mesh* m, unsigned dim, char const* name,
    unsigned ncomps, double const* data);
/* size-field producers (element_gradients.cpp, recover_by_volume.cpp,
 * size_from_hessian.cpp, size.cpp:90-135); weights may be null (all 1) */
void orc_element_gradients(orc_mesh* m, char const* name);
void orc_recover_by_volume(orc_mesh* m, char const* name);
void orc_size_from_hessian(orc_mesh* m, char const* hess_name, double const* weights,
    double min_h, double max_h);
void orc_identity_size(orc_mesh* m, char const* name);
/* Householder least-squares fit of coarsen_fit (4 to 16 rows of 4) */
unsigned orc_qr_fit(unsigned npts, double const* a, double const* b, double* x);
/* a field with an enum osh_transfer type (0 not, 1 pointwise, 2 conserve) */
void orc_add_field_tt(orc_mesh* m, unsigned dim, char const* name,
    unsigned ncomps, double const* data, int transfer);
void orc_free_field(orc_mesh* m, unsigned dim, char const* name);

double orc_min_quality(orc_mesh* m);
void orc_qualities(orc_mesh* m, double* out);
void orc_measure_edges(orc_mesh* m, double* out);
void orc_find_indset(orc_mesh* m, unsigned ent_dim,
    unsigned const* candidates, double const* goodness, unsigned* out);

unsigned orc_refine_by_size(orc_mesh* m, double qual_floor);
void orc_uniformly_refine(orc_mesh* m);
unsigned orc_coarsen_by_size(orc_mesh* m, double qual_floor, double size_floor);
unsigned orc_coarsen_slivers(orc_mesh* m, double qual_floor, unsigned nlayers);
unsigned orc_swap_slivers(orc_mesh* m, double good_qual, unsigned nlayers);
int orc_adapt(orc_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops);
int orc_adapt_timed(orc_mesh* m, double size_floor, double good_qual,
    unsigned nlayers, unsigned max_ops, double* stats);
unsigned orc_warp_to_limit(orc_mesh* m, double qual_floor);

/* pure kernels, for known-answer tests */
double orc_tet_quality(double const* x12);
double orc_triangle_quality(double const* x9);
double orc_triangle_xy_quality(double const* x9);
double orc_doubles_min(double const* a, unsigned n);
double orc_doubles_max(double const* a, unsigned n);
unsigned orc_choose_edge_swap(unsigned ring_size, double const* edge_x6,
    double const* ring_x, double* quality);
unsigned orc_swap_nint_edges(unsigned ring_size);
void orc_swap_int_edges(unsigned ring_size, unsigned code, unsigned* out);
void orc_invert_map(unsigned nin, unsigned const* in, unsigned nout,
    unsigned* out, unsigned* offsets);

/* anisotropic extension: metric from a vertex Hessian field (6 comps) */
void orc_eigen3(double const* h6, double* lam, double* V9);
double orc_metric_quality(unsigned dim, double const* x, double const* m6, int signed_z);
double orc_metric_length(double const* a6, double const* b6, double const* x0,
    double const* x1);
void orc_metric_from_hessian(orc_mesh* m, char const* hessian, double eps,
    double hmin, double hmax, char const* out_name);
/* the candidate-quality loops of refine (edges) and coarsen, for tests:
 * cands / codes updated in place, goodness out (ne, or 2 ne values) */
void orc_refine_qualities(orc_mesh* m, unsigned* cands, double qual_floor,
    int require_better, double* out);
void orc_coarsen_qualities(orc_mesh* m, unsigned* codes, double qual_floor,
    int require_better, double* out);

/* round limit of the independent set (default 100, as indset.cpp:33) */
void orc_set_indset_max_rounds(unsigned n);

/* verbosity of the per-pass prints (reference prints them from rank 0) */
void orc_set_verbose(int v);

#ifdef __cplusplus
}
#endif

#endif
