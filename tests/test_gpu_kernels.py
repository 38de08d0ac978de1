"""Kernel-level C ABI (include/ohx.h) against independent host answers:
numpy scans and stable sorts, the oracle's quality and independent set."""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """Minimal hipMalloc/hipMemcpy helper for passing device pointers."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        self.ptrs.append(p)
        return p

    def empty(self, n, dtype):
        return self.put(np.zeros(n, dtype=dtype))

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.fixture(scope="module")
def lib():
    l = ML.load("product")
    if l is None:
        pytest.fail("product library not built")
    d = l.dll
    d.ohx_exscan_u32.restype = C.c_uint
    d.ohx_exscan_u32.argtypes = [C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    d.ohx_up_from_down.restype = C.c_uint
    d.ohx_up_from_down.argtypes = [C.c_uint] * 4 + [C.c_void_p] * 5
    d.ohx_element_qualities_f64.restype = C.c_uint
    d.ohx_element_qualities_f64.argtypes = [C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    d.ohx_min_quality_f64.restype = C.c_double
    d.ohx_min_quality_f64.argtypes = [C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    d.ohx_find_indset.restype = C.c_uint
    d.ohx_find_indset.argtypes = [C.c_uint] + [C.c_void_p] * 6
    d.ohx_last_error.restype = C.c_char_p
    return d


@pytest.fixture
def dev():
    d = Dev()
    yield d
    d.free()


@pytest.mark.parametrize("n", [0, 1, 2, 4095, 4096, 4097, 12289, 1 << 20, 30_000_001])
def test_exscan(lib, dev, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 7, n, dtype=np.uint32)
    pa, po = dev.put(a), dev.empty(n + 1, np.uint32)
    assert lib.ohx_exscan_u32(pa, n, po, None) == 0, lib.ohx_last_error()
    got = dev.get(po, n + 1, np.uint32)
    want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)]).astype(np.uint32)
    assert np.array_equal(got, want)


def test_exscan_repeated_epochs(lib, dev):
    """Many consecutive multi-tile scans reuse the tile-state words."""
    rng = np.random.default_rng(5)
    for k in range(40):
        n = int(rng.integers(5000, 200000))
        a = rng.integers(0, 3, n, dtype=np.uint32)
        pa, po = dev.put(a), dev.empty(n + 1, np.uint32)
        assert lib.ohx_exscan_u32(pa, n, po, None) == 0
        got = dev.get(po, n + 1, np.uint32)
        assert got[-1] == a.sum() and np.array_equal(got[1:], np.cumsum(a).astype(np.uint32))


@pytest.mark.parametrize("hd,ld,nlows,nhighs", [(3, 0, 5000, 20000), (2, 1, 30000, 20000), (3, 2, 9000, 4000),
                                                 (1, 0, 100000, 300000)])
def test_up_from_down_is_stable_counting_sort(lib, dev, hd, ld, nlows, nhighs):
    deg = [[1, 0, 0, 0], [2, 1, 0, 0], [3, 3, 1, 0], [4, 6, 4, 1]][hd][ld]
    rng = np.random.default_rng(hd * 10 + ld)
    lows = rng.integers(0, nlows, nhighs * deg, dtype=np.uint32)
    p = dev.put(lows)
    po, pa, pd = dev.empty(nlows + 1, np.uint32), dev.empty(len(lows), np.uint32), dev.empty(len(lows), np.uint32)
    assert lib.ohx_up_from_down(hd, ld, nhighs, nlows, p, po, pa, pd, None) == 0
    order = np.argsort(lows, kind="stable")  # uses by low, then use index
    want_offs = np.concatenate([[0], np.cumsum(np.bincount(lows, minlength=nlows))]).astype(np.uint32)
    assert np.array_equal(dev.get(po, nlows + 1, np.uint32), want_offs)
    assert np.array_equal(dev.get(pa, len(lows), np.uint32), (order // deg).astype(np.uint32))
    assert np.array_equal(dev.get(pd, len(lows), np.uint32), (order % deg).astype(np.uint32))


@pytest.mark.parametrize("dim,n", [(2, 64), (3, 12)])
def test_element_quality_kernel_matches_oracle(lib, dev, dim, n):
    x, c = ML.box(dim, n)
    rng = np.random.default_rng(dim)
    x = x + rng.normal(0, 0.2 / n, x.shape)
    if dim == 2:
        x[:, 2] = 0
    orc = ML.Mesh(ML.load("oracle"), dim, x, c)
    want = orc.qualities()
    px, pc, pq = dev.put(x.ravel()), dev.put(c.ravel()), dev.empty(len(c), np.float64)
    assert lib.ohx_element_qualities_f64(dim, len(c), pc, px, pq, None) == 0
    got = dev.get(pq, len(c), np.float64)
    assert got.tobytes() == want.tobytes()
    mq = lib.ohx_min_quality_f64(dim, len(c), pc, px, None)
    assert mq == orc.min_quality() == want[1:].min()


def test_find_indset_matches_oracle(lib, dev):
    """Independent set over the edge star of a 3D mesh, random goodness with
    many exact ties."""
    x, c = ML.box(3, 6)
    orc = ML.Mesh(ML.load("oracle"), 3, x, c)
    orc.derive_model_full()
    ne = orc.count(1)
    rng = np.random.default_rng(11)
    cand = (rng.random(ne) < 0.6).astype(np.uint32)
    good = np.round(rng.random(ne), 2)  # ties on purpose
    want = orc.find_indset(1, cand, good)
    offs, adj = orc.star(1, 3)
    po, pa = dev.put(offs), dev.put(adj)
    pc, pg, ps = dev.put(cand), dev.put(good), dev.empty(ne, np.uint32)
    r = lib.ohx_find_indset(ne, po, pa, pc, pg, ps, None)
    assert r != 0xFFFFFFFF, lib.ohx_last_error()
    got = dev.get(ps, ne, np.uint32)
    assert np.array_equal(got, want)
    # independence and maximality within the filter
    src = np.repeat(np.arange(ne), np.diff(offs))
    assert not np.any(got[src] & got[adj])
    blocked = np.zeros(ne, bool)
    np.logical_or.at(blocked, src, got[adj] == 1)
    assert np.all(got[(cand == 1) & ~blocked] == 1)
