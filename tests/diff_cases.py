"""Planted differences between pairs of meshes for the mesh_diff tests.

Each case is (name, piece_a, piece_b, tol, floor, superset), built from the
reference-written fixtures of tests/golden/vtu/ read by the VTU oracle, so
the pairs are plain data: the tests write both pieces to files and give them
to the reference's mesh_diff, the oracle's and the product's.
"""
import copy
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import vtu as V  # noqa: E402

GOLD = os.path.join(HERE, "golden", "vtu")


def piece(name):
    with open(os.path.join(GOLD, name + ".vtu"), "rb") as f:
        return V.read_vtu(f.read())


def _set_tag(p, d, name, arr, ncomps=1):
    lst = [t for t in p.tags.get(d, []) if t[0] != name]
    lst.append((name, ncomps, arr))
    p.tags[d] = lst


def _edit_tag(p, d, name, fn):
    for k, (n, c, a) in enumerate(p.tags[d]):
        if n == name:
            a = a.copy()
            fn(a)
            p.tags[d][k] = (n, c, a)
            return
    raise KeyError(name)


def _drop_tag(p, d, name):
    p.tags[d] = [t for t in p.tags[d] if t[0] != name]


def _drop_edges(p):
    del p.conn[1]
    p.tags.pop(1, None)


def cases():
    base3 = piece("adapt3d_n2")
    base2 = piece("warp2d_sine_n16")
    out = []

    def pair(name, a, b, tol=1e-6, floor=0.0, superset=False):
        out.append((name, a, b, tol, floor, superset))

    pair("same", base3, copy.deepcopy(base3))
    b = copy.deepcopy(base3)
    x = [t for t in base3.tags[0] if t[0] == "coordinates"][0][2]
    k = int(np.nonzero(x[1:])[0][3]) + 1
    _edit_tag(b, 0, "coordinates", lambda y: y.__setitem__(k, y[k] * (1 + 1e-3)))
    pair("coordinate_beyond_tolerance", base3, b)
    pair("coordinate_within_tolerance", base3, b, tol=1e-1)
    b = copy.deepcopy(base3)
    _edit_tag(b, 0, "coordinates", lambda x: x.__setitem__(0, 0.25))
    pair("entry_zero_not_looked_at", base3, b)
    # two entries with the same value moved alike: both printed as the max pair
    b = copy.deepcopy(base3)
    same = np.nonzero(x == 0.5)[0][1:3]
    _edit_tag(b, 0, "coordinates", lambda y: y.__setitem__(same, 0.5 * (1 + 1e-4)))
    pair("two_max_pairs", base3, b)
    a, b = copy.deepcopy(base3), copy.deepcopy(base3)
    n = a.nverts
    _set_tag(a, 0, "tiny", np.full(n, 1e-12))
    _set_tag(b, 0, "tiny", np.full(n, 2e-12))
    pair("below_floor", a, b, floor=1e-10)
    pair("above_zero_floor", a, b)
    a, b = copy.deepcopy(base3), copy.deepcopy(base3)
    _set_tag(a, 0, "zeros", np.zeros(n))
    _set_tag(b, 0, "zeros", np.zeros(n))
    pair("zeros_are_nan_and_ignored", a, b)
    b = copy.deepcopy(base3)
    _drop_tag(b, 0, "adapt_size")
    pair("tag_only_on_first", base3, b)
    pair("tag_only_on_first_superset", base3, b, superset=True)
    pair("tag_only_on_second", b, base3)
    pair("tag_only_on_second_superset", b, base3, superset=True)
    a, b = copy.deepcopy(base3), copy.deepcopy(base3)
    _set_tag(a, 3, "t", np.arange(a.count(3), dtype=np.float64))
    _set_tag(b, 3, "t", np.arange(2 * b.count(3), dtype=np.float64), ncomps=2)
    pair("component_counts", a, b)
    a, b = copy.deepcopy(base3), copy.deepcopy(base3)
    _set_tag(a, 3, "t", np.arange(a.count(3), dtype=np.uint32))
    _set_tag(b, 3, "t", np.arange(b.count(3), dtype=np.float64))
    pair("types", a, b)
    b = copy.deepcopy(base3)
    _edit_tag(b, 1, "class_id", lambda y: y.__setitem__(3, y[3] + 1))
    pair("integer_contents", base3, b)
    b = copy.deepcopy(base3)
    c = b.conn[3].copy()
    c[[0, 1]] = c[[1, 0]]
    b.conn[3] = c
    pair("connectivity", base3, b)
    pair("counts", base3, piece("metric3d_n2"))
    pair("dimensions", base2, base3)
    b = copy.deepcopy(base2)
    _drop_edges(b)
    pair("only_first_has_edges", base2, b)
    pair("only_second_has_edges", b, base2)
    pair("only_second_has_edges_superset", b, base2, superset=True)
    return out
