"""Batches: many meshes adapted in one launch sequence must each end exactly
as mesh_adapt leaves them alone.

Every component of the product's batch run is extracted and compared, array
by array, with the oracle's standalone run of the same input mesh (the
oracle has no batching: it adapts one mesh per call, as the reference
does), and the per-component return values must match.
"""
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


def _aniso_mesh(lib, spec):
    x, c = ML.box(3, spec["n"])
    m = ML.Mesh(lib, 3, x * spec["side"], c)
    m.derive_model_full()
    m.add_field(0, "metric", W.constant_metric(m.count(0), spec["rot"], spec["lengths"]).ravel())
    return m


def _iso2d_mesh(lib, n, h, warp):
    x, c = ML.box(2, n)
    m = ML.Mesh(lib, 2, x, c)
    m.derive_model_full()
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), h))
    if warp:
        xs = m.coords().reshape(-1, 3)
        m.add_field(0, "warp", W.sine_warp_2d(xs).ravel())
        m.warp_to_limit(0.2)
        m.free_field(0, "warp")
    return m


def _check(batch, singles, max_ops):
    r, res, stats = batch.batch_adapt(1.0 / 3.0, 0.3, 4, max_ops)
    assert r == int((res < 0).sum())
    changed = 0
    for c, single in enumerate(singles):
        want = single.adapt(1.0 / 3.0, 0.3, 4, max_ops)
        assert int(res[c]) == want, (c, int(res[c]), want)
        got = batch.batch_extract(c)
        d = P.diff_states(got, single)
        assert d is None, "component %d: %s" % (c, d)
        changed += want != 0
    assert changed > 0
    if changed:
        assert stats[0] > 0 and stats[1] > 0
    return res


def test_batch_3d_anisotropic_components_equal_single_runs(product, oracle):
    specs = W.batch_specs(6, seed=7, n_lo=2, n_hi=4, h_lo=0.05, h_hi=0.5)
    parts = [_aniso_mesh(product, s) for s in specs]
    batch = ML.Mesh.batch(product, parts)
    assert batch.batch_size() == len(specs)
    # the untouched batch splits back into its inputs
    for c, p in enumerate(parts):
        assert P.diff_states(batch.batch_extract(c), p) is None
    singles = [_aniso_mesh(oracle, s) for s in specs]
    _check(batch, singles, 50)


def test_batch_2d_isotropic_components_equal_single_runs(product, oracle):
    cases = [(6, 0.08, True), (10, 0.05, False), (4, 0.3, True), (8, 0.02, True)]
    batch = ML.Mesh.batch(product, [_iso2d_mesh(product, *c) for c in cases])
    singles = [_iso2d_mesh(oracle, *c) for c in cases]
    _check(batch, singles, 50)


def test_batch_operation_limit_is_per_component(product, oracle):
    # with this seed and limit two components reach the limit, two finish
    specs = W.batch_specs(4, seed=12, n_lo=2, n_hi=3, h_lo=0.03, h_hi=0.3)
    batch = ML.Mesh.batch(product, [_aniso_mesh(product, s) for s in specs])
    singles = [_aniso_mesh(oracle, s) for s in specs]
    res = _check(batch, singles, 5)
    assert (res < 0).any() and (res > 0).any()


@pytest.mark.parametrize("handoff", [0.0, 1.0])
def test_batch_schedule_extremes_change_nothing(product, oracle, handoff):
    """All batch passes (never handing over) and immediate hand-over to
    single-mesh runs both give the single runs' meshes."""
    import ctypes as C
    d = product.dll
    d.ohx_set_batch_handoff.argtypes = [C.c_double]
    d.ohx_set_batch_handoff(handoff)
    try:
        specs = W.batch_specs(5, seed=5, n_lo=2, n_hi=4, h_lo=0.05, h_hi=0.5)
        batch = ML.Mesh.batch(product, [_aniso_mesh(product, s) for s in specs])
        singles = [_aniso_mesh(oracle, s) for s in specs]
        _check(batch, singles, 50)
    finally:
        d.ohx_set_batch_handoff(0.25)


def test_retiring_finished_components_changes_nothing(product, oracle):
    """With retirement forced at every step the batch still equals the
    single runs (finished components leave the live batch and come back at
    the end, in order)."""
    import ctypes as C
    d = product.dll
    d.ohx_set_batch_retire_min.argtypes = [C.c_double]
    d.ohx_set_batch_retire_min(0.0)
    d.ohx_set_batch_handoff.argtypes = [C.c_double]
    d.ohx_set_batch_handoff(0.0)
    try:
        specs = W.batch_specs(5, seed=3, n_lo=2, n_hi=4, h_lo=0.05, h_hi=0.5)
        batch = ML.Mesh.batch(product, [_aniso_mesh(product, s) for s in specs])
        singles = [_aniso_mesh(oracle, s) for s in specs]
        _check(batch, singles, 50)
        cases = [(6, 0.08, True), (4, 0.3, True), (8, 0.02, True)]
        batch = ML.Mesh.batch(product, [_iso2d_mesh(product, *c) for c in cases])
        singles = [_iso2d_mesh(oracle, *c) for c in cases]
        _check(batch, singles, 50)
    finally:
        d.ohx_set_batch_retire_min(1e6)
        d.ohx_set_batch_handoff(0.25)


@pytest.mark.parametrize("nthreads", [1, 4])
def test_adapt_many_concurrent_streams_equal_single_runs(product, oracle, nthreads):
    """Independent meshes adapted concurrently by worker threads, each on a
    stream of its own, end as their single runs (3D anisotropic and 2D
    isotropic meshes in one call)."""
    specs = W.batch_specs(6, seed=9, n_lo=2, n_hi=4, h_lo=0.05, h_hi=0.5)
    cases = [(6, 0.08, True), (10, 0.05, False), (8, 0.02, True)]
    meshes = [_aniso_mesh(product, s) for s in specs] + [_iso2d_mesh(product, *c) for c in cases]
    singles = [_aniso_mesh(oracle, s) for s in specs] + [_iso2d_mesh(oracle, *c) for c in cases]
    r, res, stats = ML.Mesh.adapt_many(product, meshes, max_ops=50, nthreads=nthreads)
    assert r == int((res < 0).sum())
    for i, (m, s) in enumerate(zip(meshes, singles)):
        assert int(res[i]) == s.adapt(1.0 / 3.0, 0.3, 4, 50), i
        d = P.diff_states(m, s)
        assert d is None, "mesh %d: %s" % (i, d)
    assert stats[0] > 0 and stats[1] > 0 and stats[3] > 0
