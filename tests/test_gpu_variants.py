"""Every tuned variant of a kernel gives the same result.

The online tuner picks among interchangeable variants of some kernels (lanes
per edge, plain or wave-combined counting, hash insertion, scan tiles and
look-back window); which one runs depends on timings, so each is pinned here
in turn and the adapt runs of the parity cases are compared with the
oracle's after every step.
"""
import ctypes as C

import pytest

import meshlib as ML
import parity_cases as P

pytestmark = pytest.mark.gpu

CASES = {k: v for k, v in dict(P.CASES, **P.ANISO_CASES).items()
         if k in ("warp2d_sine_n16", "warp3d_rot_n4", "aniso2d_tanh_n16", "aniso3d_rot_n6")}
PINS = [
    ({"coarsen_quality_2d_f64": 1, "coarsen_quality_3d_f64": 1}, "collapse qualities, one lane per edge"),
    ({"coarsen_quality_2d_f64": 8, "coarsen_quality_3d_f64": 8}, "collapse qualities, eight lanes per edge"),
    ({"upfd_count_%d_%d" % (h, l): 1 for h in (1, 2, 3) for l in range(h)}, "up_from_down, wave-combined counting"),
    ({"upfd_count_%d_%d" % (h, l): 0 for h in (1, 2, 3) for l in range(h)}, "up_from_down, plain counting"),
    ({"upfd_rank_split_%d_%d" % (h, l): 208 for h in (1, 2, 3) for l in range(h)},
     "up_from_down ranks, buckets held in 8 lanes (longer ones read from memory)"),
    ({"upfd_rank_split_%d_%d" % (h, l): 264 for h in (1, 2, 3) for l in range(h)},
     "up_from_down ranks, buckets held in 64 lanes"),
    ({"upfd_rank_split_%d_%d" % (h, l): 4 for h in (1, 2, 3) for l in range(h)},
     "up_from_down ranks, 4 lanes reading memory"),
    ({"lows_hash_insert_1": 1, "lows_hash_insert_2": 1}, "hash insertion, home slots first"),
    ({"lows_hash_insert_1": 0, "lows_hash_insert_2": 0}, "hash insertion, compare-and-swap only"),
    ({"scan": 44}, "scans, 256-tile look-back window"),
    ({"scan": 16}, "scans, 16 items per thread"),
]


@pytest.fixture(scope="module")
def product():
    lib = ML.load("product")
    if lib is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    lib.dll.ohx_tuning_force.argtypes = [C.c_char_p, C.c_int]
    return lib


@pytest.fixture(scope="module")
def oracle():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle library not built: make -C oracle")
    return lib


@pytest.mark.parametrize("pin", range(len(PINS)), ids=[p[1] for p in PINS])
@pytest.mark.parametrize("case", sorted(CASES))
def test_pinned_variant_matches_oracle(case, pin, product, oracle):
    pins = PINS[pin][0]
    d = product.dll
    for name, v in pins.items():
        d.ohx_tuning_force(name.encode(), v)
    try:
        n = 0
        for (la, ma), (lb, mb) in zip(CASES[case](product, True), CASES[case](oracle, False)):
            assert la == lb
            diff = P.diff_states(ma, mb)
            assert diff is None, "%s after %s: %s" % (case, la, diff)
            n += 1
        assert n >= 2
    finally:
        for name in pins:
            d.ohx_tuning_force(name.encode(), -1)
