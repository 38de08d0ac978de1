"""The C-ABI library loads and exports every symbol the headers declare.
No compute calls (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import meshlib as ML

INCLUDE = os.path.join(ML.ROOT, "include")


def declared():
    names = set()
    for h in ("omega_h.h", "ohx.h"):
        src = open(os.path.join(INCLUDE, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b((?:osh|ohx)_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_api():
    names = declared()
    ref_api = ["osh_init", "osh_fini", "osh_free", "osh_read_vtk", "osh_write_vtk", "osh_new",
               "osh_build_ents", "osh_dim", "osh_nelems", "osh_nverts", "osh_count", "osh_down",
               "osh_up", "osh_up_offs", "osh_up_dirs", "osh_star", "osh_star_offs", "osh_coords",
               "osh_own_rank", "osh_own_id", "osh_global", "osh_new_field", "osh_get_field",
               "osh_free_field", "osh_nfields", "osh_field", "osh_components", "osh_new_label",
               "osh_get_label", "osh_free_label", "osh_new_global", "osh_accumulate_to_owner",
               "osh_conform", "osh_mark_classified", "osh_ghost", "osh_num_ghost_layers",
               "osh_adapt", "osh_identity_size"]
    assert len(ref_api) == 38
    missing = [n for n in ref_api if n not in names]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    if not os.path.exists(ML.PRODUCT_LIB):
        pytest.fail("product library not built: run __graft_entry__.build()")
    out = subprocess.run(["nm", "-D", "--defined-only", ML.PRODUCT_LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(declared() - exported)
    assert not missing, missing


def test_library_loads_without_a_device():
    # loading must not touch the GPU; every entry point needs one at call time
    lib = C.CDLL(ML.PRODUCT_LIB)
    assert hasattr(lib, "osh_adapt") and hasattr(lib, "ohx_mesh_adapt")


def test_oracle_and_product_seams_match():
    prod = subprocess.run(["nm", "-D", "--defined-only", ML.PRODUCT_LIB], capture_output=True,
                          text=True, check=True).stdout
    orc = subprocess.run(["nm", "-D", "--defined-only", ML.ORACLE_LIB], capture_output=True,
                         text=True, check=True).stdout
    p = {l.split()[-1][len("ohx_mesh_"):] for l in prod.splitlines() if " ohx_mesh_" in l}
    o = {l.split()[-1][len("orc_"):] for l in orc.splitlines() if " orc_" in l}
    for name in ML._SIG:
        assert ML._ALIAS["ohx_mesh_"].get(name, name) in p, "product seam lacks " + name
        assert name in o, "oracle lacks " + name
