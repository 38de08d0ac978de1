"""The VTU oracle (oracle/vtu.py) against files the compiled reference wrote
(tests/golden/vtu/, made by tests/golden/make_vtu_golden.py).  CPU only.

For every fixture mesh, rebuilt in the oracle's C library from the same
script: the oracle's writer reproduces the reference's file byte for byte
(uncompressed, zlib-compressed and the .pvtu index) and the oracle's reader
returns exactly the mesh's arrays.
"""
import os
import sys

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P

sys.path.insert(0, os.path.join(ML.ROOT, "oracle"))
import vtu as V  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtu")


def _gold(name):
    with open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    return lib


def assert_same_piece(a, b):
    assert (a.dim, a.full, a.nverts) == (b.dim, b.full, b.nverts)
    assert sorted(a.conn) == sorted(b.conn)
    for d in a.conn:
        assert np.array_equal(a.conn[d], b.conn[d]), d
    for d in range(a.dim + 1):
        # a reader adds the coordinates first, whatever their place in the
        # writer's mesh: compare the tags by name
        ta = {n: (c, x) for n, c, x in a.tags.get(d, [])}
        tb = {n: (c, x) for n, c, x in b.tags.get(d, [])}
        assert sorted(ta) == sorted(tb), d
        for n in ta:
            (ca, x), (cb, y) = ta[n], tb[n]
            assert ca == cb and x.dtype == y.dtype, (d, n)
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (d, n)


@pytest.mark.parametrize("case", P.VTU_CASES)
def test_oracle_writes_the_reference_bytes(case, oracle):
    p = V.piece_of(P.vtu_case(case, oracle))
    assert V.write_vtu(p) == _gold(case + ".vtu")
    assert V.write_vtu(p, compress=True) == _gold(case + "_z.vtu")


@pytest.mark.parametrize("case", P.VTU_CASES)
def test_oracle_reads_the_reference_files(case, oracle):
    want = V.piece_of(P.vtu_case(case, oracle))
    assert_same_piece(V.read_vtu(_gold(case + ".vtu")), want)
    assert_same_piece(V.read_vtu(_gold(case + "_z.vtu")), want)


def test_oracle_pvtu_index(oracle):
    p = V.piece_of(P.vtu_case("warp2d_sine_n16", oracle))
    assert V.write_pvtu(p, "warp2d_sine_n16.vtu") == _gold("warp2d_sine_n16.pvtu")


def test_oracle_catches_planted_mistakes(oracle):
    """A dropped tag, a swapped connectivity entry or a changed value makes
    the bytes differ from the reference's."""
    p = V.piece_of(P.vtu_case("metric3d_n2", oracle))
    gold = _gold("metric3d_n2.vtu")
    q = V.piece_of(P.vtu_case("metric3d_n2", oracle))
    q.tags[0] = q.tags[0][:-1]
    assert V.write_vtu(q) != gold
    q = V.piece_of(P.vtu_case("metric3d_n2", oracle))
    c = q.conn[3].copy()
    c[[0, 1]] = c[[1, 0]]
    q.conn[3] = c
    assert V.write_vtu(q) != gold
    q = V.piece_of(P.vtu_case("metric3d_n2", oracle))
    name, nc, a = q.tags[0][-1]
    a = a.copy()
    a[0] = np.nextafter(a[0], 2.0)
    q.tags[0][-1] = (name, nc, a)
    assert V.write_vtu(q) != gold
    assert V.write_vtu(p) == gold


def test_reference_reads_oracle_files(oracle, tmp_path):
    """Where the reference checker is built: it reads the oracle's files
    back into the same mesh (uncompressed and compressed)."""
    ref, refz = ML.load("ref"), ML.load("ref_zlib")
    if ref is None or refz is None:
        pytest.skip("reference checker not built here (oracle/_ref)")
    for case in P.VTU_CASES:
        m = P.vtu_case(case, oracle)
        p = V.piece_of(m)
        for lib, comp in ((ref, False), (refz, True)):
            path = str(tmp_path / ("%s_%d.vtu" % (case, comp)))
            with open(path, "wb") as f:
                f.write(V.write_vtu(p, compress=comp))
            got = ML.Mesh.read_vtk(lib, path)
            assert P.diff_states(got, m) is None, (case, comp)


def _ref_diff(lib, a, b, tol, floor, superset, tmp_path, capfd):
    import ctypes as C
    pa, pb = str(tmp_path / "a.vtu"), str(tmp_path / "b.vtu")
    for path, p in ((pa, a), (pb, b)):
        with open(path, "wb") as f:
            f.write(V.write_vtu(p))
    ma, mb = ML.Mesh.read_vtk(lib, pa), ML.Mesh.read_vtk(lib, pb)
    f = lib.dll.ref_mesh_diff
    f.restype, f.argtypes = C.c_uint, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_uint]
    C.CDLL(None).fflush(None)  # what earlier calls left in C's stdout buffer
    capfd.readouterr()
    r = f(ma.h, mb.h, tol, floor, 1 if superset else 0)
    return int(r), capfd.readouterr().out.splitlines()


def test_mesh_diff_oracle_matches_the_reference(tmp_path, capfd):
    """Verdict and printed lines of the oracle's mesh_diff equal the compiled
    reference's on every planted difference (where the reference is built)."""
    import diff_cases as D
    ref = ML.load("ref")
    if ref is None:
        pytest.skip("reference checker not built here (oracle/_ref)")
    seen = set()
    for name, a, b, tol, floor, superset in D.cases():
        want = _ref_diff(ref, a, b, tol, floor, superset, tmp_path, capfd)
        got = V.mesh_diff(a, b, tol, floor, superset)
        assert got == want, name
        seen.add(got[0])
    assert seen == {0, 1}


def test_mesh_diff_oracle_verdicts():
    """The verdicts the planted differences are built to give (this holds
    without the reference too)."""
    import diff_cases as D
    expect0 = {"same", "coordinate_within_tolerance", "entry_zero_not_looked_at", "below_floor",
               "zeros_are_nan_and_ignored", "tag_only_on_first_superset", "tag_only_on_second_superset",
               "only_second_has_edges_superset"}
    for name, a, b, tol, floor, superset in D.cases():
        r, lines = V.mesh_diff(a, b, tol, floor, superset)
        assert r == (0 if name in expect0 else 1), (name, lines)
        if name == "two_max_pairs":
            assert sum("max difference pair" in ln for ln in lines) == 2
