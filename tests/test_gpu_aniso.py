"""Anisotropic extension on the GPU against the oracle's restatement of it:
bit-exact FP64 (same operation order, no fused multiply-adds)."""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P
from test_gpu_kernels import Dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def product():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built")
    return lib


@pytest.fixture(scope="module")
def oracle():
    return ML.load("oracle")


@pytest.mark.parametrize("case", sorted(P.ANISO_CASES))
def test_aniso_workload_parity(case, product, oracle):
    n = 0
    for (la, ma), (lb, mb) in zip(P.ANISO_CASES[case](product, True), P.ANISO_CASES[case](oracle)):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None), (case, la)
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)
        n += 1
    assert n >= 5


def test_eigen_and_metric_kernels_bitwise(product, oracle):
    rng = np.random.default_rng(9)
    n = 4096
    b = rng.normal(size=(n, 3, 3)) * (10.0 ** rng.uniform(-2, 3, (n, 1, 1)))
    h = b + np.transpose(b, (0, 2, 1))
    h6 = np.stack([h[:, 0, 0], h[:, 0, 1], h[:, 0, 2], h[:, 1, 1], h[:, 1, 2], h[:, 2, 2]], 1).copy()
    d = product.dll
    d.ohx_eigen3_f64.restype = C.c_uint
    d.ohx_eigen3_f64.argtypes = [C.c_uint] + [C.c_void_p] * 4
    d.ohx_metric_from_hessian_f64.restype = C.c_uint
    d.ohx_metric_from_hessian_f64.argtypes = [C.c_uint, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                              C.c_void_p, C.c_void_p]
    dev = Dev()
    try:
        ph = dev.put(h6.ravel())
        pl, pv, pm = dev.empty(3 * n, np.float64), dev.empty(9 * n, np.float64), dev.empty(6 * n, np.float64)
        assert d.ohx_eigen3_f64(n, ph, pl, pv, None) == 0
        assert d.ohx_metric_from_hessian_f64(n, ph, 0.01, 1e-3, 0.1, pm, None) == 0
        lam, V, M = dev.get(pl, 3 * n, np.float64), dev.get(pv, 9 * n, np.float64), dev.get(pm, 6 * n, np.float64)
    finally:
        dev.free()
    o = oracle.dll
    o.orc_eigen3.argtypes = [C.c_void_p] * 3
    for k in range(0, n, 7):
        l2, v2 = np.empty(3), np.empty(9)
        hk = np.ascontiguousarray(h6[k])
        o.orc_eigen3(hk.ctypes.data, l2.ctypes.data, v2.ctypes.data)
        assert lam[3 * k:3 * k + 3].tobytes() == l2.tobytes()
        assert V[9 * k:9 * k + 9].tobytes() == v2.tobytes()
    # metric against the oracle through a one-vertex mesh per tensor
    x, c = ML.box(3, 1)
    for k in range(0, 64):
        m = ML.Mesh(oracle, 3, x, c)
        m.add_field(0, "hessian", np.tile(h6[k], (8, 1)).ravel())
        m.metric_from_hessian("hessian", 0.01, 1e-3, 0.1)
        want = m.field(0, "metric")[:6]
        assert M[6 * k:6 * k + 6].tobytes() == want.tobytes()
