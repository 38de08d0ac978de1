"""The anisotropic-metric extension of the oracle, pinned by mathematics
(there is no reference for it): eigendecomposition against LAPACK,
invariances of the metric quality, closed forms of metric lengths."""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML

P = C.c_void_p


@pytest.fixture(scope="module")
def orc():
    d = ML.load("oracle").dll
    d.orc_eigen3.argtypes = [P, P, P]
    d.orc_metric_quality.restype = C.c_double
    d.orc_metric_quality.argtypes = [C.c_uint, P, P, C.c_int]
    d.orc_metric_length.restype = C.c_double
    d.orc_metric_length.argtypes = [P, P, P, P]
    d.orc_tet_quality.restype = C.c_double
    d.orc_tet_quality.argtypes = [P]
    d.orc_triangle_xy_quality.restype = C.c_double
    d.orc_triangle_xy_quality.argtypes = [P]
    return d


def sym6(a):
    return np.array([a[0, 0], a[0, 1], a[0, 2], a[1, 1], a[1, 2], a[2, 2]])


def full(m6):
    return np.array([[m6[0], m6[1], m6[2]], [m6[1], m6[3], m6[4]], [m6[2], m6[4], m6[5]]])


def ptr(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(P)


def test_jacobi_eigen_against_lapack(orc):
    rng = np.random.default_rng(3)
    for k in range(500):
        b = rng.normal(size=(3, 3)) * (10.0 ** rng.uniform(-3, 3))
        h = b + b.T
        if k % 5 == 0:
            h = np.outer(rng.normal(size=3), rng.normal(size=3))
            h = h + h.T  # rank deficient
        h6 = sym6(h)
        lam = np.empty(3)
        V = np.empty(9)
        orc.orc_eigen3(ptr(h6), lam.ctypes.data_as(P), V.ctypes.data_as(P))
        V = V.reshape(3, 3)
        scale = np.abs(h).max()
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-13)
        assert np.allclose(V @ np.diag(lam) @ V.T, h, atol=1e-13 * scale)
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(h), atol=1e-13 * scale)


def test_metric_quality_reduces_to_v1_for_isotropic_metric(orc):
    rng = np.random.default_rng(4)
    for _ in range(200):
        x = rng.normal(size=(4, 3))
        h = 10.0 ** rng.uniform(-3, 1)
        m = np.tile(sym6(np.eye(3) / h ** 2), (4, 1))
        q1 = orc.orc_tet_quality(ptr(x))
        qm = orc.orc_metric_quality(3, ptr(x), ptr(m), 1)
        assert abs(qm - q1) <= 1e-12 * max(1.0, abs(q1))


def test_metric_quality_affine_invariance(orc):
    """Quality in metric M of points x equals the plain quality of the
    points mapped by any factor L with L^T L = M (here a random A)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        x = rng.normal(size=(4, 3))
        a = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        if np.linalg.det(a) < 0:
            a[0] = -a[0]
        m = np.tile(sym6(a.T @ a), (4, 1))
        qm = orc.orc_metric_quality(3, ptr(x), ptr(m), 1)
        q = orc.orc_tet_quality(ptr(x @ a.T))
        assert abs(qm - q) <= 1e-10 * max(1.0, abs(q))


def test_metric_quality_regular_in_metric_is_one(orc):
    rng = np.random.default_rng(6)
    reg = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    for _ in range(50):
        a = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        m = np.tile(sym6(a.T @ a), (4, 1))
        x = reg @ np.linalg.inv(a).T  # maps to the regular tet under a
        qm = orc.orc_metric_quality(3, ptr(x), ptr(m), 1)
        assert abs(abs(qm) - 1.0) < 1e-12


def test_metric_length_closed_form(orc):
    rng = np.random.default_rng(7)
    for _ in range(200):
        b = rng.normal(size=(3, 3))
        ma, mb = b @ b.T + np.eye(3), np.diag(rng.uniform(0.5, 2, 3))
        x0, x1 = rng.normal(size=3), rng.normal(size=3)
        e = x1 - x0
        want = np.sqrt(e @ ((ma + mb) / 2) @ e)
        got = orc.orc_metric_length(ptr(sym6(ma)), ptr(sym6(mb)), ptr(x0), ptr(x1))
        assert abs(got - want) <= 1e-13 * want


def test_metric_from_hessian_clamps_sizes():
    import parity_cases as P_
    from omega_h_v1_amd import workloads as W
    orc = ML.load("oracle")
    x, c = ML.box(3, 4)
    m = ML.Mesh(orc, 3, x, c)
    h = W.tanh_front_hessian(x, a=10.0)
    m.add_field(0, "hessian", h.ravel())
    m.metric_from_hessian("hessian", 0.8, 0.04, 0.3)
    mt = m.field(0, "metric").reshape(-1, 6)
    for k in range(len(x)):
        ev = np.linalg.eigvalsh(full(mt[k]))
        sizes = 1 / np.sqrt(ev)
        assert sizes.min() >= 0.04 * (1 - 1e-12) and sizes.max() <= 0.3 * (1 + 1e-12)
        # the front normal (1,1,1)/sqrt(3) is an eigenvector
        n = np.ones(3) / np.sqrt(3)
        mn = full(mt[k]) @ n
        assert np.allclose(mn, (n @ mn) * n, atol=1e-9 * np.abs(mt[k]).max())
