"""FP32-mixed precision: every adapt decision must equal the FP64 path's.

The product runs with FP32 screening of the split and collapse qualities
(ohx_set_precision(1)) and is compared, after every step, with the oracle's
FP64 run: the meshes must be bit-identical, which is the decision identity
the mixed mode promises.  The largest |q32 - q64| seen where both were
computed must stay well inside the screening margin (1e-4), and the screen
must actually have settled some decisions in FP32.
"""
import ctypes as C

import pytest

import meshlib as ML
import parity_cases as P

pytestmark = pytest.mark.gpu

MARGIN = 1e-4
# the FP32 variants pinned on: the screen (FP32 rejects, FP64 for the rest)
# or the certified pass (FP32 settles every decision it can certify, FP64
# only near a threshold or a competing goodness); lanes per edge alternate
MODES = {
    "screen": {"coarsen_quality_2d_mixed": 18, "coarsen_quality_3d_mixed": 24,
               "refine_quality_2d_mixed": 1, "refine_quality_3d_mixed": 1, "swap_quality_3d_mixed": 1},
    "certified": {"coarsen_quality_2d_mixed": 36, "coarsen_quality_3d_mixed": 34,
                  "refine_quality_2d_mixed": 1, "refine_quality_3d_mixed": 1, "swap_quality_3d_mixed": 1},
}


@pytest.fixture(scope="module")
def product():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    d = lib.dll
    d.ohx_set_precision.argtypes = [C.c_int]
    d.ohx_mixed_stats.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
    d.ohx_tuning_force.argtypes = [C.c_char_p, C.c_int]
    d.ohx_set_precision(1)
    yield lib
    for name in MODES["screen"]:
        d.ohx_tuning_force(name.encode(), -1)
    d.ohx_set_precision(0)


def _pin(lib, mode):
    # the FP32 variants are tuned choices: pin one on, so that every
    # decision here goes through it
    for name, v in MODES[mode].items():
        lib.dll.ohx_tuning_force(name.encode(), v)


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    return lib


def _stats(lib):
    e, a, b = C.c_double(), C.c_ulonglong(), C.c_ulonglong()
    lib.dll.ohx_mixed_stats(C.byref(e), C.byref(a), C.byref(b))
    return e.value, a.value, b.value


CASES = dict(P.CASES)
CASES.update(P.ANISO_CASES)
CASES["c2_n256"] = lambda lib, dev=False: P.c2_steps(lib, on_device=dev)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("case", sorted(CASES))
def test_mixed_decisions_equal_fp64(case, mode, product, oracle):
    _pin(product, mode)
    product.dll.ohx_mixed_stats_reset()
    n = 0
    for (la, ma), (lb, mb) in zip(CASES[case](product, True), CASES[case](oracle, False)):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None), (case, la)
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)
        n += 1
    assert n >= 2
    err, screened, reeval = _stats(product)
    print("%s: max |q32 - q64| = %.3g, settled in FP32 %d, re-evaluated in FP64 %d" % (case, err, screened, reeval))
    assert err < MARGIN / 10
    assert screened + reeval > 0


def test_mixed_stats_account_for_every_candidate(product):
    """On C1 the screen sees candidates, and the counters reset."""
    _pin(product, "screen")
    product.dll.ohx_mixed_stats_reset()
    assert _stats(product) == (0.0, 0, 0)
    for _ in P.c1_steps(product, True):
        pass
    err, screened, reeval = _stats(product)
    assert screened + reeval > 0
