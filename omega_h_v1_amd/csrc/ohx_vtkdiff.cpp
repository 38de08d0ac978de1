// vtkdiff: compares two VTU meshes with mesh_diff, as the reference's
// test_vtkdiff.cpp:14-96 does -- same options (-tolerance, -Floor,
// -superset, -help), same defaults (1e-6, 0), exit status 2 when the
// meshes differ, 0 when they agree.  Both meshes are read into HBM and
// compared there (ohx_mesh_read_vtk, ohx_mesh_diff).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ohx.h"

int main(int argc, char** argv) {
  double tol = 1e-6, floor = 0;
  bool get_tol = false, get_floor = false, help = false;
  unsigned superset = 0;
  char const* fa = nullptr;
  char const* fb = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (get_tol) {
      tol = std::atof(argv[i]);
      get_tol = false;
    } else if (get_floor) {
      floor = std::atof(argv[i]);
      get_floor = false;
    } else if (!std::strcmp(argv[i], "-tolerance")) {
      get_tol = true;
    } else if (!std::strcmp(argv[i], "-Floor")) {
      get_floor = true;
    } else if (!std::strcmp(argv[i], "-help")) {
      help = true;
    } else if (!std::strcmp(argv[i], "-superset")) {
      superset = 1;
    } else if (!fa) {
      fa = argv[i];
    } else if (!fb) {
      fb = argv[i];
    }
  }
  if (get_tol) {
    std::printf("-tolerance needs an argument\n");
    return -1;
  }
  if (get_floor) {
    std::printf("-Floor needs an argument\n");
    return -1;
  }
  if (help || !fa || !fb) {
    std::printf("\nusage: %s [options] file1.vtu file2.vtu\n", argv[0]);
    std::printf("    or:  %s [-help]             (usage)\n\n", argv[0]);
    std::printf("    -help (Print this summary and exit.)\n");
    std::printf("    -tolerance <$val> (Overrides the default tolerance of 1.0E-6.)\n");
    std::printf("    -Floor <$val> (Overrides the default floor tolerance of 0.0.)\n");
    std::printf("    -superset (The second mesh may hold extra tags and dimensions.)\n");
    return 0;
  }
  ohx_mesh* a = ohx_mesh_read_vtk(fa);
  ohx_mesh* b = a ? ohx_mesh_read_vtk(fb) : nullptr;
  if (!a || !b) {
    std::fprintf(stderr, "vtkdiff: %s\n", ohx_last_error());
    if (a) ohx_mesh_free(a);
    return 1;
  }
  unsigned differ = ohx_mesh_diff(a, b, tol, floor, superset);
  ohx_mesh_free(a);
  ohx_mesh_free(b);
  if (differ == OHX_ERROR) {
    std::fprintf(stderr, "vtkdiff: %s\n", ohx_last_error());
    return 1;
  }
  std::fflush(stdout);
  return differ ? 2 : 0;
}
