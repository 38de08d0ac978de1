"""Builds omega_h_v1_amd/lib/libomega_h_amd.so with hipcc for gfx950, and
the command-line tools of omega_h_v1_amd/bin/ (vtkdiff) against it.

Each .hip source compiles to an object in _build/ (in parallel), then one
shared library is linked.  Flags that matter for results:

* -ffp-contract=off: no fused multiply-add anywhere, so FP64 results round
  exactly like the reference's x86-64 SSE2 build;
* -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 math.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libomega_h_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["ohx_runtime.hip", "ohx_prims.hip", "ohx_mesh.hip", "ohx_adapt.hip", "ohx_transfer.hip", "ohx_fields.hip", "ohx_batch.hip", "ohx_vtk.hip",
           "ohx_diff.hip", "ohx_api.hip", "ohx_noalloc.hip"]
# host-only command-line tools linked against the library
TOOLS = {"vtkdiff": "ohx_vtkdiff.cpp"}
BIN = os.path.join(HERE, "bin")

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, warn in results:
            if warn:
                print(warn)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    for name, src in TOOLS.items():
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, src),
               "-o", os.path.join(BIN, name), "-L" + os.path.dirname(LIB), "-lomega_h_amd",
               "-Wl,-rpath,$ORIGIN/../lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("building %s failed:\n%s" % (name, r.stderr[-4000:]))
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
