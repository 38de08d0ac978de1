"""Adjacency derivations on the GPU against the oracle: every star and the
element dual, on reduced meshes (where the general star and the dual from
vertices run) and full ones, including adapted meshes of uneven valence."""
import numpy as np
import pytest

import meshlib as ML
import parity_cases as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def product():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    return lib


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    return lib


def _adapted(oracle, case):
    m = None
    for _, m in dict(P.CASES, **P.SIZE_CASES)[case](oracle):
        pass
    return m.dim, m.coords().reshape(-1, 3).copy(), m.down(m.dim, 0).reshape(-1, m.dim + 1).copy()


def _meshes(oracle):
    out = []
    for dim, n in ((2, 7), (3, 4)):
        x, c = ML.box(dim, n)
        out.append(("box%dd" % dim, dim, x, c))
    for case in ("warp3d_rot_n4", "dye_warp2d"):
        dim, x, c = _adapted(oracle, case)
        out.append((case, dim, x, c))
    return out


def _compare(a, b, what):
    assert a[0].shape == b[0].shape and np.array_equal(a[0], b[0]), what + " offsets"
    assert np.array_equal(a[1], b[1]), what + " adjacency"


def test_reduced_mesh_stars_and_dual(product, oracle):
    for name, dim, x, c in _meshes(oracle):
        mp, mo = ML.Mesh(product, dim, x, c), ML.Mesh(oracle, dim, x, c)
        # the general star (no edges derived yet) and the dual from vertices
        _compare(mp.star(0, dim), mo.star(0, dim), "%s star(0,%d)" % (name, dim))
        assert np.array_equal(mp.dual(), mo.dual()), name + " dual"


def test_full_mesh_stars_and_dual(product, oracle):
    for name, dim, x, c in _meshes(oracle):
        mp, mo = ML.Mesh(product, dim, x, c), ML.Mesh(oracle, dim, x, c)
        mp.derive_model_full()
        mo.derive_model_full()
        for l in range(dim):
            for h in range(l + 1, dim + 1):
                _compare(mp.star(l, h), mo.star(l, h), "%s star(%d,%d)" % (name, l, h))
        assert np.array_equal(mp.dual(), mo.dual()), name + " dual"


def _maps_match_fresh(m, what):
    """The mesh's side maps and upward adjacencies equal the ones a copy
    with only the entities' vertices derives (asking for a map the mesh
    lacks derives it, so later passes carry it too)."""
    fresh = m.clone()
    downs = ((3, 1), (3, 2), (2, 1)) if m.dim == 3 else ((2, 1),)
    ups = ((0, 1), (0, 3), (1, 3)) if m.dim == 3 else ((0, 1), (0, 2))
    for h, l in downs:
        assert np.array_equal(m.down(h, l), fresh.down(h, l)), "%s down(%d,%d)" % (what, h, l)
    for l, h in ups:
        for a, b, part in zip(m.up(l, h), fresh.up(l, h), ("offsets", "adjacency", "directions")):
            assert np.array_equal(a, b), "%s up(%d,%d) %s" % (what, l, h, part)
    assert np.array_equal(m.dual(), fresh.dual()), what + " dual"


def test_swapped_mesh_maps_match_derivation(product):
    """A swap pass hands the next mesh its tets' edges, the edges' tets and
    the element dual built from the old mesh's maps instead of deriving them
    again; on the warped 4^3 box, after every swap pass that changes the
    mesh, they equal the derived ones."""
    swaps = 0
    for label, m in P.CASES["warp3d_rot_n4"](product):
        if not label.startswith("warp"):
            continue
        for _ in range(4):
            if not m.swap_slivers(0.3, 4):
                break
            swaps += 1
            _maps_match_fresh(m, "%s, swap %d" % (label, swaps))
    assert swaps >= 3


@pytest.mark.parametrize("case", ["warp3d_rot_n4", "warp2d_sine_n16"])
def test_coarsened_mesh_maps_match_derivation(product, case):
    """A collapse pass carries the side maps and upward adjacencies it read
    to the next mesh (survivors renumbered, products from their old entity,
    fused sides to the side they merge into); after every sliver collapse
    pass that changes the mesh, and after every mesh_adapt of the case,
    they equal the derived ones."""
    passes = 0
    for label, m in P.CASES[case](product):
        if label.startswith("adapt"):
            _maps_match_fresh(m, label)
            continue
        if not label.startswith("warp"):
            continue
        for _ in range(3):
            if not m.coarsen_slivers(0.3, 4):
                break
            passes += 1
            _maps_match_fresh(m, "%s, collapse %d" % (label, passes))
    assert passes >= 1
