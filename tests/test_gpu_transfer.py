"""Solution transfer on the GPU: element fields tagged conserve or
pointwise follow refine, coarsen and swap exactly as in the oracle (which
equals the compiled reference step for step, tests/test_oracle_vs_ref.py).
"""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P
from omega_h_v1_amd import workloads as W

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


@pytest.mark.parametrize("case", sorted(P.TRANSFER_CASES))
def test_transfer_equals_oracle_every_step(case, product, oracle):
    n = 0
    for (la, ma), (lb, mb) in zip(P.TRANSFER_CASES[case](product), P.TRANSFER_CASES[case](oracle)):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None), la
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)
        assert ma.tag_names(ma.dim) == mb.tag_names(mb.dim), la
        n += 1
    assert n >= 4


def test_pointwise_fields_need_3d_for_coarsening(product):
    """The reference asserts a 3D mesh in coarsen_fit; the product reports
    the same condition as an error."""
    m = P.make_mesh(product, 2, 6)
    nel = m.count(2)
    m.add_field(2, "p", np.ones(nel), transfer=P.POINTWISE)
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), 0.5))
    with pytest.raises(RuntimeError, match="3D"):
        m.coarsen_by_size(0.2, 1.0 / 3.0)


def test_conserved_total_at_size(product):
    """A 2D warp-and-adapt on 2 x 128^2 triangles: the conserved total stays
    within rounding of its start (a size-independent check)."""
    m = P.make_mesh(product, 2, 128, on_device=True)
    x = m.coords().reshape(-1, 3)
    ev = m.down(2, 0).reshape(-1, 3).astype(np.int64)
    cen = x[ev].sum(1) / 3
    mass = 1.0 + cen[:, 0] ** 2
    m.add_field(2, "mass", mass, transfer=P.CONSERVE)
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), 1.0 / 100.0))
    m.add_field(0, "warp", W.sine_warp_2d(x).ravel())
    m.warp_to_limit(0.2)
    assert m.adapt(1.0 / 3.0, 0.3, 4, 50) != 0
    got = m.tag(2, "mass", f64=True)
    assert len(got) == m.count(2)
    assert abs(got.sum() - mass.sum()) <= 1e-12 * abs(mass.sum())


def test_adapt_many_carries_fields(product, oracle):
    """Meshes with transfer fields adapted concurrently on worker streams
    end as the oracle's single runs."""
    specs = [(3, 3, 0.2), (3, 4, 0.15), (2, 8, 0.08), (3, 2, 0.3)]

    def build(lib, dim, n, h):
        m = P.make_mesh(lib, dim, n)
        P.add_element_fields(m)
        m.add_field(0, "adapt_size", W.uniform_size(m.count(0), h))
        return m
    meshes = [build(product, *s) for s in specs]
    singles = [build(oracle, *s) for s in specs]
    r, res, stats = ML.Mesh.adapt_many(product, meshes, max_ops=50, nthreads=4)
    for i, (m, s) in enumerate(zip(meshes, singles)):
        assert int(res[i]) == s.adapt(1.0 / 3.0, 0.3, 4, 50), i
        d = P.diff_states(m, s)
        assert d is None, "mesh %d: %s" % (i, d)
