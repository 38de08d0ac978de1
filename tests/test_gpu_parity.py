"""Parity of the HIP product against the oracle (serial C restatement).

Bar: bit-exact connectivity, tags and coordinates (integer and FP64 alike:
the FP64 path reproduces the reference's x86-64 arithmetic exactly), after
every pass of every scripted workload.  Every call goes through the
product's C ABI (include/ohx.h).
"""
import ctypes as C

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


def _run_pair(case, product, oracle, on_device=False):
    gp = P.CASES[case](product, on_device)
    go = P.CASES[case](oracle, False)
    steps = 0
    for (la, ma), (lb, mb) in zip(gp, go):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None), (case, la)
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)
        steps += 1
    return steps


def test_fp64_sqrt_div_bitwise(product):
    """Device sqrt and division must round exactly like the host."""
    rng = np.random.default_rng(20261202)
    n = 1 << 22
    a = np.concatenate([
        rng.random(n // 4),
        np.exp(rng.uniform(-700, 700, n // 4)),
        rng.random(n // 4) * 1e-300,
        (rng.integers(0, 1 << 62, n // 4, dtype=np.int64)).view(np.float64),
    ])
    a = np.abs(a[np.isfinite(a)])
    b = np.exp(rng.uniform(-300, 300, len(a)))
    special = np.array([0.0, 1.0, 2.0, 4.0, 0.25, 3.0, 1e-310, 5e-324, 1.7976931348623157e308])
    a = np.concatenate([special, a])
    b = np.concatenate([np.full(len(special), 3.0), b])
    import ctypes
    lib = product.dll
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = a.nbytes
    ptrs = []
    for _ in range(4):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        ptrs.append(p)
    H2D, D2H = 1, 2
    assert hip.hipMemcpy(ptrs[0], a.ctypes.data_as(C.c_void_p), ctypes.c_size_t(nbytes), H2D) == 0
    assert hip.hipMemcpy(ptrs[1], b.ctypes.data_as(C.c_void_p), ctypes.c_size_t(nbytes), H2D) == 0
    f = lib.ohx_fp64_selftest
    f.restype = C.c_uint
    f.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    assert f(len(a), ptrs[0], ptrs[1], ptrs[2], ptrs[3], None) == 0
    lib.ohx_synchronize()
    s = np.empty_like(a)
    q = np.empty_like(a)
    assert hip.hipMemcpy(s.ctypes.data_as(C.c_void_p), ptrs[2], ctypes.c_size_t(nbytes), D2H) == 0
    assert hip.hipMemcpy(q.ctypes.data_as(C.c_void_p), ptrs[3], ctypes.c_size_t(nbytes), D2H) == 0
    for p in ptrs:
        hip.hipFree(p)
    bad_s = np.nonzero(s.view(np.uint64) != np.sqrt(a).view(np.uint64))[0]
    bad_q = np.nonzero(q.view(np.uint64) != (a / b).view(np.uint64))[0]
    assert len(bad_s) == 0, "sqrt differs at %d inputs, e.g. %r" % (len(bad_s), a[bad_s[:3]])
    assert len(bad_q) == 0, "div differs at %d inputs, e.g. %r / %r" % (len(bad_q), a[bad_q[:3]], b[bad_q[:3]])


def test_c1_parity(product, oracle):
    assert _run_pair("c1", product, oracle) == 4


def test_c1_device_generator_parity(product, oracle):
    """The device Kuhn generator builds the same mesh as the host one."""
    assert _run_pair("c1", product, oracle, on_device=True) == 4


@pytest.mark.parametrize("case", ["warp2d_sine_n16", "warp2d_rot_n32", "warp3d_rot_n4"])
def test_warp_workload_parity(case, product, oracle):
    assert _run_pair(case, product, oracle) > 4


def test_c2_full_size_parity(product, oracle):
    """The 2D drop-in configuration at full size (131072 triangles)."""
    gp = P.c2_steps(product, on_device=True)
    go = P.c2_steps(oracle)
    for (la, ma), (lb, mb) in zip(gp, go):
        assert ma.last == mb.last
        d = P.diff_states(ma, mb)
        assert d is None, "c2 after %s: %s" % (la, d)
