"""ohx_exscan_u32_noalloc, the scan entry point that allocates nothing,
against ohx_exscan_u32 (the existing entry point, which takes its tile
words from the library's arena): bit-identical outputs on two calls in a
row, the second with a smaller input (the fixed scratch keeps the first
call's contents), and a clear error for an input beyond the scratch's
maximum.  Everything goes through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
MAX_ITEMS = 1 << 27


@pytest.fixture(scope="module")
def api():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    d = lib.dll
    for name in ("ohx_exscan_u32", "ohx_exscan_u32_noalloc"):
        f = getattr(d, name)
        f.restype = C.c_uint
        f.argtypes = [C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    d.ohx_last_error.restype = C.c_char_p
    hip = C.CDLL("libamdhip64.so")
    return d, hip


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 4))) == 0
    return p


def _scan_both(d, hip, a):
    n = len(a)
    pa, po1, po2 = _dev(hip, a.nbytes), _dev(hip, 4 * (n + 1)), _dev(hip, 4 * (n + 1))
    try:
        if n:
            assert hip.hipMemcpy(pa, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        assert d.ohx_exscan_u32(pa, n, po1, None) == 0, d.ohx_last_error()
        assert d.ohx_exscan_u32_noalloc(pa, n, po2, None) == 0, d.ohx_last_error()
        assert hip.hipDeviceSynchronize() == 0
        o1 = np.empty(n + 1, dtype=np.uint32)
        o2 = np.empty(n + 1, dtype=np.uint32)
        assert hip.hipMemcpy(o1.ctypes.data_as(C.c_void_p), po1, C.c_size_t(o1.nbytes), D2H) == 0
        assert hip.hipMemcpy(o2.ctypes.data_as(C.c_void_p), po2, C.c_size_t(o2.nbytes), D2H) == 0
    finally:
        for p in (pa, po1, po2):
            hip.hipFree(p)
    return o1, o2


def test_noalloc_scan_matches_exscan_twice(api):
    d, hip = api
    rng = np.random.default_rng(7)
    # large first (many tiles, a multi-step look-back), then smaller: the
    # second call meets the first call's tile words in the scratch
    for n in (3_000_017, 70_001, 4096, 1, 0):
        a = rng.integers(0, 64, n, dtype=np.uint32)
        o1, o2 = _scan_both(d, hip, a)
        assert np.array_equal(o1, o2), n
        want = np.zeros(n + 1, dtype=np.uint64)
        np.cumsum(a, out=want[1:])
        assert np.array_equal(o2, want.astype(np.uint32)), n


def test_noalloc_scan_rejects_oversize(api):
    d, hip = api
    d.ohx_clear_error()
    r = d.ohx_exscan_u32_noalloc(None, MAX_ITEMS + 1, None, None)
    assert r == 0xFFFFFFFF
    msg = d.ohx_last_error()
    assert msg is not None and b"2^27" in msg
    d.ohx_clear_error()
    # still works afterwards
    a = np.arange(1000, dtype=np.uint32)
    o1, o2 = _scan_both(d, hip, a)
    assert np.array_equal(o1, o2)
