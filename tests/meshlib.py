"""Loader of the three libraries the tests compare, over the package's
ctypes front-end (omega_h_v1_amd/seam.py).  All three expose the same
mesh-level seam under different prefixes:

* ``omega_h_v1_amd/lib/libomega_h_amd.so`` (prefix ``ohx_mesh_``) -- the
  product;
* ``oracle/_build/libohv1_oracle.so`` (prefix ``orc_``) -- the serial C
  restatement used as the checker;
* ``oracle/_ref/libohv1_ref_serial.so`` / ``..._omp.so`` / ``..._zlib.so``
  (prefix ``ref_``) -- the reference itself, compiled from its sources (only
  where it was built; the zlib build writes compressed VTU arrays).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from omega_h_v1_amd.seam import Lib, Mesh, box, PRODUCT_LIB, _DOWN_DEG, _SIG, _ALIAS, _ptr  # noqa: E402,F401

ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libohv1_oracle.so")
REF_SERIAL_LIB = os.path.join(ROOT, "oracle", "_ref", "libohv1_ref_serial.so")
REF_OMP_LIB = os.path.join(ROOT, "oracle", "_ref", "libohv1_ref_omp.so")
REF_ZLIB_LIB = os.path.join(ROOT, "oracle", "_ref", "libohv1_ref_zlib.so")


def load(which):
    """which in {'product', 'oracle', 'ref', 'ref_omp', 'ref_zlib'}; None if
    not built."""
    path, prefix = {
        "product": (PRODUCT_LIB, "ohx_mesh_"),
        "oracle": (ORACLE_LIB, "orc_"),
        "ref": (REF_SERIAL_LIB, "ref_"),
        "ref_omp": (REF_OMP_LIB, "ref_"),
        "ref_zlib": (REF_ZLIB_LIB, "ref_"),
    }[which]
    if not os.path.exists(path):
        return None
    return Lib(path, prefix)
