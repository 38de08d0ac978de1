"""The drop-in boundary itself: the same sequence of public omega_h.h calls
(osh_new, osh_build_ents, osh_new_global, osh_new_field, osh_new_label,
osh_adapt, the adjacency and field accessors, osh_identity_size) made on the
product (arrays in device memory) and on the compiled reference (its serial
build: arrays in host memory) must leave identical meshes.

The reference marks osh_new meshes parallel (omega_h.cpp:177-180), so the
caller supplies global numbers (the identity, one rank); the product's mesh
is serial and takes them without effect.  Where the reference checker is not
built, the product is compared with the oracle's seam on the same mesh.
"""
import ctypes as C

import numpy as np
import pytest

import meshlib as ML
import parity_cases as P
from omega_h_v1_amd import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
POINTWISE, CONSERVE = 1, 2
_u, _p, _d, _s = C.c_uint, C.c_void_p, C.c_double, C.c_char_p

_SIG = {
    "osh_init": (None, [_p, _p]),
    "osh_new": (_p, [_u]),
    "osh_free": (None, [_p]),
    "osh_build_ents": (_p, [_p, _u, _u]),
    "osh_new_global": (_p, [_p, _u]),
    "osh_new_field": (_p, [_p, _u, _s, _u, C.c_int]),
    "osh_get_field": (_p, [_p, _u, _s]),
    "osh_new_label": (_p, [_p, _u, _s, _u]),
    "osh_get_label": (_p, [_p, _u, _s]),
    "osh_adapt": (_u, [_p, _d, _d, _u, _u]),
    "osh_dim": (_u, [_p]),
    "osh_count": (_u, [_p, _u]),
    "osh_nelems": (_u, [_p]),
    "osh_nverts": (_u, [_p]),
    "osh_down": (_p, [_p, _u, _u]),
    "osh_up": (_p, [_p, _u, _u]),
    "osh_up_offs": (_p, [_p, _u, _u]),
    "osh_up_dirs": (_p, [_p, _u, _u]),
    "osh_star": (_p, [_p, _u, _u]),
    "osh_star_offs": (_p, [_p, _u, _u]),
    "osh_coords": (_p, [_p]),
    "osh_nfields": (_u, [_p, _u]),
    "osh_field": (_s, [_p, _u, _u]),
    "osh_components": (_u, [_p, _u, _s]),
    "osh_identity_size": (None, [_p, _s]),
}
_DEG = [[1, 0, 0, 0], [2, 1, 0, 0], [3, 3, 1, 0], [4, 6, 4, 1]]


class Api:
    """One library's osh_* entry points, with the copies its memory needs."""

    def __init__(self, path, device):
        self.dll = C.CDLL(path)
        self.device = device
        for name, (res, args) in _SIG.items():
            f = getattr(self.dll, name)
            f.restype, f.argtypes = res, args
        if device:
            self.hip = C.CDLL("libamdhip64.so")
        argc, argv = C.c_int(0), C.POINTER(C.c_char_p)()
        self.dll.osh_init(C.byref(argc), C.byref(argv))

    def __getattr__(self, name):
        return getattr(self.dll, name)

    def put(self, ptr, a):
        a = np.ascontiguousarray(a)
        if not a.nbytes:
            return
        if self.device:
            assert self.hip.hipMemcpy(C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0
        else:
            C.memmove(ptr, a.ctypes.data, a.nbytes)

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        if not out.nbytes:
            return out
        if self.device:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), D2H) == 0
        else:
            C.memmove(out.ctypes.data, ptr, out.nbytes)
        return out


def _source(dim, n):
    """A full, classified box mesh built by the oracle (the caller's input)."""
    m = P.make_mesh(ML.load("oracle"), dim, n)
    ents = {d: m.down(d, 0) for d in range(1, dim + 1)}
    counts = m.counts()
    labels = {(d, name): m.tag(d, name) for d in range(dim + 1) for name in ("class_dim", "class_id")}
    return m.coords(), counts, ents, labels


def _drive(api, dim, n, h):
    """The caller's program: build, annotate, adapt, read everything back."""
    x, counts, ents, labels = _source(dim, n)
    m = api.osh_new(dim)
    api.osh_build_ents(m, 0, counts[0])
    for d in range(1, dim + 1):
        api.put(api.osh_build_ents(m, d, counts[d]), ents[d])
    for d in range(dim + 1):
        api.put(api.osh_new_global(m, d), np.arange(counts[d], dtype=np.uint64))
    api.put(api.osh_new_field(m, 0, b"coordinates", 3, POINTWISE), x)
    for (d, name), a in labels.items():
        api.put(api.osh_new_label(m, d, name.encode(), 1), a)
    api.put(api.osh_new_field(m, 0, b"adapt_size", 1, POINTWISE), W.uniform_size(counts[0], h))
    # a conserved element field rides along
    api.put(api.osh_new_field(m, dim, b"mass", 1, CONSERVE), 1.0 + np.arange(counts[dim]) % 7)
    ret = api.osh_adapt(m, 1.0 / 3.0, 0.3, 4, 50)
    api.osh_identity_size(m, b"ident")
    out = {"ret": ret, "dim": api.osh_dim(m), "nelems": api.osh_nelems(m), "nverts": api.osh_nverts(m)}
    cnt = [api.osh_count(m, d) for d in range(dim + 1)]
    out["counts"] = cnt
    for d in range(1, dim + 1):
        out["down%d" % d] = api.get(api.osh_down(m, d, 0), cnt[d] * (d + 1), np.uint32)
    out["down%d%d" % (dim, dim - 1)] = api.get(api.osh_down(m, dim, dim - 1), cnt[dim] * _DEG[dim][dim - 1],
                                               np.uint32)
    offs = api.get(api.osh_up_offs(m, 0, dim), cnt[0] + 1, np.uint32)
    out["up_offs"] = offs
    out["up"] = api.get(api.osh_up(m, 0, dim), int(offs[-1]), np.uint32)
    out["up_dirs"] = api.get(api.osh_up_dirs(m, 0, dim), int(offs[-1]), np.uint32)
    soffs = api.get(api.osh_star_offs(m, 0, 1), cnt[0] + 1, np.uint32)
    out["star_offs"] = soffs
    out["star"] = api.get(api.osh_star(m, 0, 1), int(soffs[-1]), np.uint32)
    out["coords"] = api.get(api.osh_coords(m), cnt[0] * 3, np.float64)
    fields = {}
    for d in range(dim + 1):
        for i in range(api.osh_nfields(m, d)):
            name = api.osh_field(m, d, i)
            nc = api.osh_components(m, d, name)
            fields[(d, name.decode())] = api.get(api.osh_get_field(m, d, name), cnt[d] * nc, np.float64)
        for name in ("class_dim", "class_id"):
            out["label%d%s" % (d, name)] = api.get(api.osh_get_label(m, d, name.encode()), cnt[d], np.uint32)
    out["fields"] = fields
    api.osh_free(m)
    return out


@pytest.fixture(scope="module")
def product():
    if ML.load("product") is None:
        pytest.fail("product library not built: run __graft_entry__.build() first")
    return Api(ML.PRODUCT_LIB, device=True)


@pytest.mark.parametrize("dim,n,h", [(2, 8, 0.05), (3, 3, 0.15)])
def test_public_api_matches_the_reference(product, dim, n, h):
    ref = ML.load("ref")
    if ref is None:
        pytest.skip("reference checker not built here (oracle/_ref)")
    got = _drive(product, dim, n, h)
    want = _drive(Api(ML.REF_SERIAL_LIB, device=False), dim, n, h)
    assert got["ret"] == want["ret"] == 1
    for k in ("dim", "nelems", "nverts", "counts"):
        assert got[k] == want[k], k
    for k, v in want.items():
        if isinstance(v, np.ndarray):
            assert v.dtype == got[k].dtype and np.array_equal(v.view(np.uint8), got[k].view(np.uint8)), k
    assert sorted(got["fields"]) == sorted(want["fields"])
    for k, v in want["fields"].items():
        assert np.array_equal(v.view(np.uint64), got["fields"][k].view(np.uint64)), k


@pytest.mark.parametrize("dim,n,h", [(2, 8, 0.05), (3, 3, 0.15)])
def test_public_api_matches_the_oracle(product, dim, n, h):
    """The same program against the oracle's seam (needs no reference)."""
    got = _drive(product, dim, n, h)
    orc = ML.load("oracle")
    m = P.make_mesh(orc, dim, n)
    m.add_field(dim, "mass", 1.0 + np.arange(m.count(dim)) % 7, transfer=CONSERVE)
    m.add_field(0, "adapt_size", W.uniform_size(m.count(0), h))
    assert m.adapt(1.0 / 3.0, 0.3, 4, 50) == got["ret"]
    assert m.counts() == got["counts"]
    for d in range(1, dim + 1):
        assert np.array_equal(m.down(d, 0), got["down%d" % d]), d
    assert np.array_equal(m.coords().view(np.uint64), got["coords"].view(np.uint64))
    assert np.array_equal(m.tag(dim, "mass", f64=True).view(np.uint64), got["fields"][(dim, "mass")].view(np.uint64))
