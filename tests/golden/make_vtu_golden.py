"""Writes the VTU fixtures of tests/golden/vtu/ with the compiled reference.

Run in the build container after `make -C oracle ref` (oracle/_ref holds the
reference built from its own sources: the serial build writes uncompressed
arrays, the USE_ZLIB=1 build compressed ones; neither is shipped).  Each
fixture is a mesh of tests/parity_cases.py:vtu_case as write_mesh_vtk
(vtk_io.cpp:805-830) writes it:

* <case>.vtu    -- uncompressed (reference built without zlib);
* <case>_z.vtu  -- zlib-compressed arrays (reference built with zlib);
* <case>.pvtu   -- the .pvtu index written for the name <case>.pvtu (its
                   piece is <case>.vtu, the same bytes as above).
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, TESTS)
sys.path.insert(0, os.path.dirname(TESTS))

import meshlib as ML  # noqa: E402
import parity_cases as P  # noqa: E402

OUT = os.path.join(HERE, "vtu")
PVTU_CASES = ["warp2d_sine_n16"]


def main():
    ref, refz = ML.load("ref"), ML.load("ref_zlib")
    if ref is None or refz is None:
        sys.exit("build the reference checkers first: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    for case in P.VTU_CASES:
        P.vtu_case(case, ref).write_vtk(os.path.join(OUT, case + ".vtu"))
        P.vtu_case(case, refz).write_vtk(os.path.join(OUT, case + "_z.vtu"))
    for case in PVTU_CASES:
        P.vtu_case(case, ref).write_vtk(os.path.join(OUT, case + ".pvtu"))
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
