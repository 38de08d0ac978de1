"""The oracle's solution transfer (element fields tagged conserve or
pointwise) on the CPU: the least-squares fit against the reference's own
QR and against numpy, and the invariants the transfer must keep.  The
step-by-step equality with the compiled reference is in
tests/test_oracle_vs_ref.py (TRANSFER_CASES)."""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    f = lib.dll.orc_qr_fit
    f.restype, f.argtypes = C.c_uint, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _fit(lib, prefix, a, b):
    f = getattr(lib.dll, prefix + "qr_fit")
    f.restype, f.argtypes = C.c_uint, [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.full(4, np.nan)
    r = f(len(a), a.ctypes.data, b.ctypes.data, x.ctypes.data)
    return int(r), x


def _systems(seed=3):
    """Fit systems as coarsen_fit builds them (a column of ones, then 3D
    points), full rank and rank-deficient (coplanar, collinear, repeated
    points)."""
    rng = np.random.default_rng(seed)
    out = []
    for npts in range(4, 17):
        for kind in ("random", "coplanar", "collinear", "repeated"):
            x = rng.uniform(-1, 1, (npts, 3))
            if kind == "coplanar":
                x[:, 2] = 0.3 * x[:, 0] - 0.2 * x[:, 1] + 0.1
            elif kind == "collinear":
                t = rng.uniform(-1, 1, npts)
                x = np.stack([t, 2 * t + 1, -t], 1)
            elif kind == "repeated":
                x[:] = x[0]
            a = np.concatenate([np.ones((npts, 1)), x], 1)
            b = rng.normal(size=npts)
            out.append((npts, kind, a, b))
    return out


def test_fit_matches_numpy_least_squares(oracle):
    for npts, kind, a, b in _systems():
        rank, x = _fit(oracle, "orc_", a, b)
        if kind == "random":
            assert rank == 4
            want = np.linalg.lstsq(a, b, rcond=None)[0]
            assert np.allclose(x, want, rtol=1e-9, atol=1e-9), (npts, x, want)
        else:
            assert rank < 4, (npts, kind)


def test_fit_reproduces_linear_data(oracle):
    rng = np.random.default_rng(5)
    for npts in range(4, 17):
        a = np.concatenate([np.ones((npts, 1)), rng.uniform(-1, 1, (npts, 3))], 1)
        c = np.array([0.5, -2.0, 3.0, 0.25])
        rank, x = _fit(oracle, "orc_", a, a @ c)
        assert rank == 4 and np.allclose(x, c, rtol=0, atol=1e-12)


def test_fit_equals_the_reference_bit_for_bit(oracle):
    ref = ML.load("ref")
    if ref is None:
        pytest.skip("reference checker not built here (oracle/_ref)")
    for npts, kind, a, b in _systems(seed=11):
        ro, xo = _fit(oracle, "orc_", a, b)
        rr, xr = _fit(ref, "ref_", a, b)
        assert ro == rr, (npts, kind)
        if ro == 4:
            assert np.array_equal(xo.view(np.uint64), xr.view(np.uint64)), (npts, kind)


@pytest.mark.parametrize("case", sorted(P.TRANSFER_CASES))
def test_conserved_totals_and_dropped_fields(case, oracle):
    """Conserved fields keep their totals through every operation (to
    rounding), pointwise fields stay defined on every element, and a field
    without a transfer type is gone after the first operation."""
    totals = None
    for label, m in P.TRANSFER_CASES[case](oracle):
        d = m.dim
        names = m.tag_names(d)
        mass = m.tag(d, "mass", f64=True)
        mom = m.tag(d, "momentum", f64=True).reshape(-1, 3)
        now = np.concatenate([[mass.sum()], mom.sum(0)])
        if totals is None:
            totals = now
        assert np.allclose(now, totals, rtol=1e-12, atol=1e-12), (label, now, totals)
        if getattr(m, "last", None):
            assert "junk" not in names
        if d == 3:
            v = m.tag(d, "velocity", f64=True)
            assert v.size == 3 * m.count(d) and np.isfinite(v).all()
