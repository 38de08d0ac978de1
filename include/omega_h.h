/* Public C API of the MI355X adapt library: a drop-in for Omega_h v1's
 * include/omega_h.h.  Names, argument meaning, enum values and return
 * values are those of the reference so existing callers relink unchanged.
 *
 * Memory: as in the reference's CUDA build, every array returned or handed
 * out here is DEVICE memory (HBM of the current HIP device), owned by the
 * mesh and invalidated by any call that changes the mesh topology
 * (osh_adapt, ...).  Errors that the reference turns into assert()/abort()
 * print a message to stderr and abort() at the same decision points.
 *
 * Each declaration cites the reference declaration it replaces.
 */
#ifndef OMEGA_H_H
#define OMEGA_H_H

#define OSH_PUBLIC __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* include/omega_h.h:10 */
typedef struct osh_mesh* osh_t;

/* include/omega_h.h:12-13: initialise the HIP runtime / release cached
 * device memory.  argc/argv are accepted and ignored (no MPI here). */
void osh_init(int* argc, char*** argv) OSH_PUBLIC;
void osh_fini(void) OSH_PUBLIC;

/* include/omega_h.h:15 */
void osh_free(osh_t m) OSH_PUBLIC;

/* include/omega_h.h:17-18: ASCII/base64 VTU of one serial mesh */
osh_t osh_read_vtk(char const* filename) OSH_PUBLIC;
void osh_write_vtk(osh_t m, char const* filename) OSH_PUBLIC;

/* include/omega_h.h:20-21: an empty full-topology mesh, and device
 * connectivity buffers for the caller to fill.  Unlike the reference (which
 * marks osh_new meshes parallel, omega_h.cpp:177-180) the mesh is serial:
 * one GPU, no MPI. */
osh_t osh_new(unsigned elem_dim) OSH_PUBLIC;
unsigned* osh_build_ents(osh_t m, unsigned ent_dim, unsigned nents) OSH_PUBLIC;

/* include/omega_h.h:23-26 */
unsigned osh_dim(osh_t m) OSH_PUBLIC;
unsigned osh_nelems(osh_t m) OSH_PUBLIC;
unsigned osh_nverts(osh_t m) OSH_PUBLIC;
unsigned osh_count(osh_t m, unsigned dim) OSH_PUBLIC;

/* include/omega_h.h:28-33: device arrays, derived lazily */
unsigned const* osh_down(osh_t m, unsigned high_dim, unsigned low_dim) OSH_PUBLIC;
unsigned const* osh_up(osh_t m, unsigned low_dim, unsigned high_dim) OSH_PUBLIC;
unsigned const* osh_up_offs(osh_t m, unsigned low_dim, unsigned high_dim) OSH_PUBLIC;
unsigned const* osh_up_dirs(osh_t m, unsigned low_dim, unsigned high_dim) OSH_PUBLIC;
unsigned const* osh_star(osh_t m, unsigned low_dim, unsigned high_dim) OSH_PUBLIC;
unsigned const* osh_star_offs(osh_t m, unsigned low_dim, unsigned high_dim) OSH_PUBLIC;

/* include/omega_h.h:35 */
double const* osh_coords(osh_t m) OSH_PUBLIC;

/* include/omega_h.h:37-39: serial ownership (rank 0, own index, global =
 * local index), device arrays */
unsigned const* osh_own_rank(osh_t m, unsigned dim) OSH_PUBLIC;
unsigned const* osh_own_id(osh_t m, unsigned dim) OSH_PUBLIC;
unsigned long const* osh_global(osh_t m, unsigned dim) OSH_PUBLIC;

/* include/omega_h.h:41-46 */
enum osh_transfer {
  OSH_TRANSFER_NOT,
  OSH_TRANSFER_POINTWISE,
  OSH_TRANSFER_CONSERVE,
  OSH_TRANSFER_INHERIT
};

/* include/omega_h.h:48-55 */
double* osh_new_field(osh_t m, unsigned dim, char const* name,
    unsigned ncomps, enum osh_transfer tt) OSH_PUBLIC;
double* osh_get_field(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;
void osh_free_field(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;
unsigned osh_nfields(osh_t om, unsigned dim) OSH_PUBLIC;
char const* osh_field(osh_t m, unsigned dim, unsigned i) OSH_PUBLIC;
unsigned osh_components(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;

/* include/omega_h.h:57-59 */
unsigned* osh_new_label(osh_t m, unsigned dim, char const* name, unsigned ncomps) OSH_PUBLIC;
unsigned* osh_get_label(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;
void osh_free_label(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;

/* include/omega_h.h:61-64: with one rank these are identities / no-ops */
unsigned long* osh_new_global(osh_t m, unsigned dim) OSH_PUBLIC;
void osh_accumulate_to_owner(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;
void osh_conform(osh_t m, unsigned dim, char const* name) OSH_PUBLIC;

/* include/omega_h.h:66-67: marked is a device array of nents words */
void osh_mark_classified(osh_t m, unsigned ent_dim,
    unsigned class_dim, unsigned class_id, unsigned* marked) OSH_PUBLIC;

/* include/omega_h.h:69-70: one GPU holds the whole mesh: no ghost layers */
void osh_ghost(osh_t m, unsigned nlayers) OSH_PUBLIC;
unsigned osh_num_ghost_layers(osh_t m) OSH_PUBLIC;

/* include/omega_h.h:72-76 (omega_h.cpp:955-1016): size_ratio_floor,
 * good_element_quality, nsliver_layers, max_passes as in the reference;
 * returns 1 if any pass changed the mesh. */
unsigned osh_adapt(osh_t m,
    double size_ratio_floor,
    double good_element_quality,
    unsigned nsliver_layers,
    unsigned max_passes) OSH_PUBLIC;

/* include/omega_h.h:78 (size.cpp:90-135) */
void osh_identity_size(osh_t m, char const* name) OSH_PUBLIC;

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif
