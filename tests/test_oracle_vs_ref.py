"""The oracle against the compiled reference itself, where it was built
(oracle/_ref, build container only).  CPU only."""
import pytest

import meshlib as ML
import parity_cases as P


@pytest.fixture(scope="module")
def libs():
    ref = ML.load("ref")
    if ref is None:
        pytest.skip("reference checker not built here (oracle/_ref)")
    return ML.load("oracle"), ref


@pytest.mark.parametrize("case", sorted(P.CASES))
def test_oracle_equals_reference(case, libs):
    orc, ref = libs
    for (la, ma), (lb, mb) in zip(P.CASES[case](orc), P.CASES[case](ref)):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None)
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)


@pytest.mark.parametrize("case", sorted(P.TRANSFER_CASES))
def test_oracle_transfer_equals_reference(case, libs):
    """Element fields of every transfer kind follow refine, coarsen and swap
    exactly as in the reference (every array, every step)."""
    orc, ref = libs
    for (la, ma), (lb, mb) in zip(P.TRANSFER_CASES[case](orc), P.TRANSFER_CASES[case](ref)):
        assert la == lb
        assert getattr(ma, "last", None) == getattr(mb, "last", None)
        d = P.diff_states(ma, mb)
        assert d is None, "%s after %s: %s" % (case, la, d)
