// mesh_diff (mesh_diff.cpp:16-180): two meshes agree when their dimension,
// per-dimension presence and counts, connectivity and tags agree; integer
// data exactly, FP64 data to a relative tolerance above an absolute floor.
// The comparisons run on the GPU over the resident arrays; only the verdict
// (and, for a failing FP64 tag, that tag) comes back to the host.  Messages
// and the 0/1 result follow the reference, including its serial maximum,
// which starts at index 1 (doubles.cpp:100-109), so the relative difference
// of entry 0 of an FP64 tag is never looked at.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "ohx_mesh.hpp"
#include "ohx_prims.hpp"

namespace ohx {

namespace {

// 1 if any a[i] != b[i]
bool u32_differ(u32 const* a, u32 const* b, size_t n) {
  if (!n) return false;
  OHX_REQUIRE(n <= 0xffffffffull, "mesh_diff: array beyond 2^32 entries");
  Buf<u32> flag(1);
  u32* fl = flag.get();
  parallel_for(1, OHX_LAMBDA(u32) { *fl = 0; });
  parallel_for((u32) n, OHX_LAMBDA(u32 i) {
    if (a[i] != b[i]) *fl = 1;  // every writer stores the same value
  });
  u32 h = 0;
  to_host(&h, fl, 1);
  return h != 0;
}

// doubles_diff (mesh_diff.cpp:16-44)
bool f64_differ(double const* a, double const* b, u32 n, double tol, double floor) {
  OHX_REQUIRE(tol >= 0 && floor >= 0, "mesh_diff: tolerance and floor must be non-negative");
  Buf<double> diffs(n);
  double* d = diffs.get();
  parallel_for(n, OHX_LAMBDA(u32 i) {
    double fa = fabs(a[i]), fb = fabs(b[i]);
    if (fa < floor && fb < floor) {
      d[i] = 0;
      return;
    }
    double fm = fb > fa ? fb : fa;
    d[i] = fabs(a[i] - b[i]) / fm;
  });
  double maxdiff = max_f64(d, n, 1);
  if (!(maxdiff > tol)) return false;
  std::printf("max relative difference %e\n", maxdiff);
  std::vector<double> hd(n), ha(n), hb(n);
  to_host(hd.data(), d, n);
  to_host(ha.data(), a, n);
  to_host(hb.data(), b, n);
  for (u32 i = 0; i < n; ++i)
    if (hd[i] == maxdiff) std::printf("max difference pair is (%e, %e), #%u\n", ha[i], hb[i], i);
  return true;
}

// tag_diff (mesh_diff.cpp:70-115)
bool tag_differ(Tag const& a, Tag const& b, u32 n, double tol, double floor) {
  if (a.ncomps != b.ncomps) {
    std::printf("tag \"%s\" sizes %u and %u differ\n", a.name.c_str(), a.ncomps, b.ncomps);
    return true;
  }
  if (a.type != b.type) {
    std::printf("tag \"%s\" types differ\n", a.name.c_str());
    return true;
  }
  size_t k = (size_t) n * a.ncomps;
  bool differ = a.type == TagType::F64 ? f64_differ(a.f.get(), b.f.get(), (u32) k, tol, floor)
                                       : u32_differ(a.u.get(), b.u.get(), k);
  if (differ) std::printf("tag \"%s\" contents differ\n", a.name.c_str());
  return differ;
}

Tag const* tag_named(Mesh& m, u32 d, std::string const& name) {
  for (auto& t : m.tags[d])
    if (t->name == name) return t.get();
  return nullptr;
}

// tags_diff (mesh_diff.cpp:117-139)
bool tags_differ(Mesh& a, Mesh& b, u32 d, u32 n, double tol, double floor, bool superset) {
  for (auto& ta : a.tags[d]) {
    Tag const* tb = tag_named(b, d, ta->name);
    if (!tb) {
      if (superset) continue;
      std::printf("tag \"%s\" only on first mesh\n", ta->name.c_str());
      return true;
    }
    if (tag_differ(*ta, *tb, n, tol, floor)) return true;
  }
  if (!superset)
    for (auto& tb : b.tags[d])
      if (!tag_named(a, d, tb->name)) {
        std::printf("tag \"%s\" only on second mesh\n", tb->name.c_str());
        return true;
      }
  return false;
}

}  // namespace

// mesh_diff (mesh_diff.cpp:141-178)
u32 mesh_diff(Mesh& a, Mesh& b, double tol, double floor, bool superset) {
  if (a.dim != b.dim) {
    std::printf("mesh dimensionalities %u and %u differ\n", a.dim, b.dim);
    return 1;
  }
  for (u32 d = 0; d <= a.dim; ++d) {
    bool ha = has_dim(a, d), hb = has_dim(b, d);
    if (!ha && !hb) continue;
    if (ha && !hb) {
      std::printf("only first mesh has dimension %u\n", d);
      return 1;
    }
    if (!ha && hb) {
      if (superset) continue;
      std::printf("only second mesh has dimension %u\n", d);
      return 1;
    }
    if (count(a, d) != count(b, d)) {
      std::printf("counts %u and %u differ for dimension %u\n", count(a, d), count(b, d), d);
      return 1;
    }
    u32 n = count(a, d);
    if (d && u32_differ(ask_down(a, d, 0), ask_down(b, d, 0), (size_t) n * (d + 1))) {
      std::printf("connectivity differs for dimension %u\n", d);
      return 1;
    }
    if (tags_differ(a, b, d, n, tol, floor, superset)) {
      std::printf("tags for dimension %u differ\n", d);
      return 1;
    }
  }
  return 0;
}

}  // namespace ohx
