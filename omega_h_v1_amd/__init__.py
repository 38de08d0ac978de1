"""MI355X-native Omega_h v1 adapt hot path.

The product is the C-ABI shared library lib/libomega_h_amd.so (HIP kernels
for gfx950 + C++ host driver), declared in include/omega_h.h and
include/ohx.h.  This package holds its sources (csrc/), the build script and
the host-side workload definitions; it has no compute of its own.
"""
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libomega_h_amd.so")


def library_path():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libomega_h_amd.so is not built: run __graft_entry__.build()")
    return LIB_PATH
