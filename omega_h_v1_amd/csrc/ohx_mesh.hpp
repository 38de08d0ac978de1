// Device-resident simplicial mesh with lazily derived adjacency.
//
// Same data model as the reference's struct mesh (mesh.cpp:21-334): element
// dimension, representation (reduced = vertices + elements only; full =
// every dimension stored), entity counts, cached down/up/star/dual
// adjacencies that every topology change discards, and named tags per
// dimension.  Every array is a Buf in HBM.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ohx_common.hpp"

namespace ohx {

enum class TagType { U32 = 1, F64 = 3 };

struct Tag {
  std::string name;
  TagType type;
  u32 ncomps;
  int transfer;  // enum osh_transfer of the public header
  Buf<u32> u;    // one of these holds the data
  Buf<double> f;
  void const* raw() const { return type == TagType::F64 ? (void const*) f.get() : (void const*) u.get(); }
};

struct Up {
  Buf<u32> offsets, adj, dirs;
};

struct Graph {
  Buf<u32> offsets, adj;
};

struct Mesh {
  u32 dim = 0;
  bool full = false;
  u32 counts[4] = {0, 0, 0, 0};
  Buf<u32> down[4][4];
  std::unique_ptr<Up> up[4][4];
  std::unique_ptr<Graph> star[4][4];
  Buf<u32> dual;
  // reflect_down's hash table of a low dimension's entities (slots of
  // entity ids keyed by their sorted vertices, mask = slots - 1), kept while
  // another reflection onto the same lows is due (3D: tets and faces onto
  // edges)
  Buf<u32> low_table[4];
  u32 low_mask[4] = {0, 0, 0, 0};
  std::vector<std::unique_ptr<Tag>> tags[4];

  Mesh(u32 d, bool f) : dim(d), full(f) {}
};

// entity bookkeeping
void set_ents(Mesh& m, u32 d, u32 n, Buf<u32>&& verts);
bool has_dim(Mesh& m, u32 d);
u32 count(Mesh& m, u32 d);
u32 const* ask_down(Mesh& m, u32 h, u32 l);
Up const& ask_up(Mesh& m, u32 l, u32 h);
Graph const& ask_star(Mesh& m, u32 l, u32 h);
u32 const* ask_dual(Mesh& m);
void overwrite(Mesh& m, Mesh&& with);

// tags
Tag* find_tag(Mesh& m, u32 d, char const* name);
Tag& add_tag_u32(Mesh& m, u32 d, char const* name, u32 ncomps, Buf<u32>&& data, int transfer = 0);
Tag& add_tag_f64(Mesh& m, u32 d, char const* name, u32 ncomps, Buf<double>&& data, int transfer = 0);
void remove_tag(Mesh& m, u32 d, char const* name);
double const* coords(Mesh& m);
// the anisotropic vertex metric (6 components), or null for the isotropic
// "adapt_size" field of the reference
double const* metric_of(Mesh& m);
// "adapt_size", or null when the mesh carries a metric
double const* size_of(Mesh& m);
u32 const* u32_tag(Mesh& m, u32 d, char const* name);

// kernels on a mesh
// nverts only feeds the traffic accounting of the kernel timers
Buf<double> element_qualities(u32 dim, u32 n, u32 const* verts_of_elems, double const* x, u32 nverts = 0,
                              double const* metric = nullptr);
Buf<double> mesh_qualities(Mesh& m);
double min_element_quality(u32 dim, u32 n, u32 const* verts_of_elems, double const* x, u32 nverts = 0,
                           double const* metric = nullptr);
double mesh_min_quality(Mesh& m);
Buf<double> measure_edges(Mesh& m);

// marks (mark.cpp)
Buf<u32> mark_down(Mesh& m, u32 hd, u32 ld, u32 const* marked_highs);
Buf<u32> mark_up(Mesh& m, u32 ld, u32 hd, u32 const* marked_lows);
void mark_dual_layers(Mesh& m, Buf<u32>& marked, u32 nlayers);
Buf<u32> mark_class(Mesh& m, u32 ent_dim, u32 target_dim, u32 target_id);
void unmark_boundary(Mesh& m, u32 ent_dim, u32* marked);
// eq: the element qualities when the caller has them; gate (device, with
// eq): nothing is marked unless *gate < good_qual
Buf<u32> mark_slivers(Mesh& m, double good_qual, u32 nlayers, double const* eq = nullptr,
                      double const* gate = nullptr);
Buf<u32> mark_part_boundary(Mesh& m);

// topology primitives (exposed through the kernel C ABI too)
void up_from_down(u32 hd, u32 ld, u32 nhighs, u32 nlows, u32 const* lows_of_highs, Up& out);
void bridge_graph(u32 n, u32 const* offs, u32 const* adj, u32* p_nedges, Buf<u32>* ev,
                  Buf<u32>* dirs);

// setup
void derive_model(Mesh& m, double crease_angle);

// fields computed on the mesh (ohx_fields.hip): element sizes; the size-field
// producers element_gradients.cpp, recover_by_volume.cpp,
// size_from_hessian.cpp; the identity size (size.cpp:90-135)
Buf<double> element_sizes(Mesh& m);
void element_gradients(Mesh& m, char const* name);
void recover_by_volume(Mesh& m, char const* name);
void size_from_hessian(Mesh& m, char const* hess_name, double const* weights, double min_h, double max_h);
void identity_size(Mesh& m, char const* name);

// the VTU format of osh_read_vtk / osh_write_vtk (zlib-compressed arrays
// when g_vtk_compress is set; read either way)
extern bool g_vtk_compress;
void write_vtk(Mesh& m, char const* path);
void read_vtk(char const* path, Mesh& out);
// mesh_diff (mesh_diff.cpp:141-178): 0 if the meshes agree, else 1 with a message
u32 mesh_diff(Mesh& a, Mesh& b, double tol, double floor, bool superset);

}  // namespace ohx
