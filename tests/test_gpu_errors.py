"""Error returns at the reference's abort points, through the product's C ABI.

The reference aborts where an adapt cannot go on:
the independent set not settling within its round limit
(indset.cpp:107-116, LOOP_NORETURN after MAX_ITERATIONS = 100 rounds) and a
split child of non-positive quality (refine_qualities.cpp:72,
assert(q > 0)).  The product's seam reports both as errors instead
(ohx_last_error), and the library works normally afterwards.  Every
kernel exits normally on these paths: the conditions are detected from
values the kernels write, never by a trap.
"""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML
from omega_h_v1_amd import workloads as W

pytestmark = pytest.mark.gpu

NOT_IN, IN, UNKNOWN = 0, 1, 2


@pytest.fixture(scope="module")
def product():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    lib.dll.ohx_set_indset_max_rounds.argtypes = [C.c_uint]
    yield lib
    lib.dll.ohx_set_indset_max_rounds(0)


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    return lib


def jacobi_rounds(offs, adj, filt, good):
    """The documented rule of indset.cpp:55-93, restated: Jacobi rounds from
    (filter ? UNKNOWN : NOT_IN_SET); returns (state, rounds to settle)."""
    n = len(filt)
    state = np.where(np.asarray(filt) != 0, UNKNOWN, NOT_IN).astype(np.uint32)
    rounds = 0
    while (state == UNKNOWN).any():
        old = state.copy()
        for i in np.nonzero(old == UNKNOWN)[0]:
            nb = adj[offs[i]:offs[i + 1]]
            if (old[nb] == IN).any():
                state[i] = NOT_IN
                continue
            wait = False
            for o in nb:
                if old[o] == NOT_IN:
                    continue
                if (good[i] == good[o] and o < i) or good[i] < good[o]:
                    wait = True
                    break
            if not wait:
                state[i] = IN
        rounds += 1
        assert rounds < 10000
    return state, rounds


def test_indset_round_limit_kernel_path(product):
    """A set whose ties form chains: with the limit one round short of what
    it needs the call fails naming indset.cpp:107-116; at exactly the rounds
    it needs it succeeds with the documented result."""
    m = ML.Mesh.box_on_device(product, 3, 3)
    m.derive_model_full()
    offs, adj = m.star(1, 3)
    ne = m.count(1)
    good = np.full(ne, 0.5)  # all equal: ties broken by index, long chains
    good[::7] = 0.25
    cand = np.ones(ne, dtype=np.uint32)
    cand[::5] = 0
    want, need = jacobi_rounds(offs, adj, cand, good)
    assert need >= 3
    d = product.dll
    try:
        d.ohx_set_indset_max_rounds(need - 1)
        with pytest.raises(RuntimeError, match="indset.cpp:107-116"):
            m.find_indset(1, cand, good)
        d.ohx_set_indset_max_rounds(need)
        got = m.find_indset(1, cand, good)
        assert np.array_equal(got, want)
    finally:
        d.ohx_set_indset_max_rounds(0)
    # the default limit (100) again, and the library still works
    assert np.array_equal(m.find_indset(1, cand, good), want)


def _c1(lib, on_device=False):
    if on_device:
        m = ML.Mesh.box_on_device(lib, 3, 8)
    else:
        x, c = ML.box(3, 8)
        m = ML.Mesh(lib, 3, x, c)
    m.derive_model_full()
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), 0.05))
    return m


def test_indset_round_limit_refine_pass(product, oracle):
    """The refine pass's first batch of rounds (run without a host check)
    must not hide the limit: with the rounds the C1 refinement set needs
    (from the oracle's qualities and the documented rule) minus one the
    pass fails, with exactly that many it gives the oracle's mesh."""
    mo = _c1(oracle)
    qf = min(mo.min_quality(), 0.3)
    d = oracle.dll
    d.orc_refine_qualities.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p]
    cand = (mo.measure_edges() > 1.0).astype(np.uint32)
    good = np.zeros(mo.count(1))
    d.orc_refine_qualities(mo.h, cand.ctypes.data_as(C.c_void_p), C.c_double(qf), 0,
                           good.ctypes.data_as(C.c_void_p))
    offs, adj = mo.star(1, 3)
    _, need = jacobi_rounds(offs, adj, cand, good)
    assert need >= 2
    mo.refine_by_size(qf)
    dp = product.dll
    try:
        dp.ohx_set_indset_max_rounds(need - 1)
        mp = _c1(product, True)
        with pytest.raises(RuntimeError, match="indset.cpp:107-116"):
            mp.refine_by_size(qf)
        dp.ohx_set_indset_max_rounds(need)
        mp = _c1(product, True)
        assert mp.refine_by_size(qf) == 1
    finally:
        dp.ohx_set_indset_max_rounds(0)
    import parity_cases as P
    assert P.diff_states(mp, mo) is None


def test_non_positive_split_quality_is_an_error(product):
    """An inverted element under a refinement: the split's child has
    non-positive quality, which the reference asserts against
    (refine_qualities.cpp:72); the product reports it as an error, and a
    fresh mesh adapts normally afterwards."""
    x, c = ML.box(3, 2)
    x = x.copy()
    centre = int(np.nonzero(np.all(np.abs(x - 0.5) < 1e-12, axis=1))[0][0])
    x[centre, 0] = 1.3  # through the x = 1 boundary: the tets on that side invert
    m = ML.Mesh(product, 3, x, c)
    m.derive_model_full()
    assert m.min_quality() < 0
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), 0.05))
    with pytest.raises(RuntimeError, match="refine_qualities.cpp:72"):
        m.refine_by_size(-1.0)
    ok = _c1(product, True)
    assert ok.refine_by_size(min(ok.min_quality(), 0.3)) == 1
