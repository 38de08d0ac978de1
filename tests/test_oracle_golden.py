"""The oracle against the fixtures written from the compiled reference
(tests/golden/make_golden.py).  CPU only."""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle not built: make -C oracle")
    return lib


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLDEN, "workload_digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def _digest(m):
    out = {}
    for k, v in m.state().items():
        out[k] = v if isinstance(v, int) else hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
    return out


def test_c1_full_arrays(oracle):
    gold = np.load(os.path.join(GOLDEN, "c1_states.npz"))
    seen = 0
    for label, m in P.c1_steps(oracle):
        st = m.state()
        keys = sorted(k.split("|", 1)[1] for k in gold.files if k.startswith(label + "|"))
        assert keys == sorted(st), label
        for k in keys:
            g = gold[label + "|" + k]
            v = np.asarray(st[k])
            assert g.dtype == v.dtype and g.shape == v.shape, (label, k)
            assert g.tobytes() == v.tobytes(), (label, k)
            seen += 1
    assert seen > 20


@pytest.mark.parametrize("case", ["c1", "warp2d_sine_n16", "warp2d_rot_n32", "warp3d_rot_n4"])
def test_workload_digests(case, oracle, digests):
    steps = digests[case]
    n = 0
    # the generator mutates one mesh: compare each step as it is produced
    for (label, m), g in zip(P.CASES[case](oracle), steps):
        n += 1
        assert label == g["label"]
        assert getattr(m, "last", None) == g["last"], label
        assert m.counts() == g["counts"], label
        assert _digest(m) == g["arrays"], label
    assert n == len(steps)


def test_c2_full_size_digest(oracle, digests):
    steps = digests["c2_n256"]
    for (label, m), g in zip(P.c2_steps(oracle), steps):
        assert label == g["label"]
        assert m.last == g["last"] and m.counts() == g["counts"]
        assert _digest(m) == g["arrays"], label


def test_swap_interior_edge_tables(oracle, known):
    d = oracle.dll
    d.orc_swap_nint_edges.restype = C.c_uint
    for rs, rows in known["swap_int_edges"].items():
        rs = int(rs)
        k = d.orc_swap_nint_edges(rs)
        assert k == rs - 3
        for code, row in enumerate(rows):
            buf = (C.c_uint * (2 * k))()
            d.orc_swap_int_edges(rs, code, buf)
            assert list(buf) == row, (rs, code)


def test_choose_edge_swap_known_answers(oracle, known):
    f = oracle.dll.orc_choose_edge_swap
    f.restype = C.c_uint
    f.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    for case in known["choose_edge_swap"]:
        e = np.array(case["edge"], dtype=np.float64)
        r = np.array(case["ring"], dtype=np.float64)
        q = C.c_double()
        code = f(case["ring_size"], e.ctypes.data, r.ctypes.data, C.byref(q))
        assert code == case["code"]
        assert q.value == case["quality"]


def test_quality_known_answers(oracle, known):
    d = oracle.dll
    for fn in ("orc_tet_quality", "orc_triangle_quality", "orc_triangle_xy_quality"):
        getattr(d, fn).restype = C.c_double
        getattr(d, fn).argtypes = [C.c_void_p]
    for name, rec in known["qualities"].items():
        x = np.array(rec["x"], dtype=np.float64)
        if "tet" in rec:
            assert d.orc_tet_quality(x.ctypes.data) == rec["tet"], name
        else:
            assert d.orc_triangle_quality(x.ctypes.data) == rec["tri"], name
            assert d.orc_triangle_xy_quality(x.ctypes.data) == rec["tri_xy"], name
