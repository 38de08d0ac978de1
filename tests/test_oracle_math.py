"""The oracle against what mathematics fixes, independent of the reference:
closed-form qualities and their invariances, the stable counting sort behind
upward adjacency, unique answers of the downward adjacency, star sets,
independent sets by a brute-force run of the documented Jacobi rule
(indset.cpp:55-117), and the best edge-swap triangulation by exhaustive
enumeration of all triangulations of the ring.  CPU only."""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import meshlib as ML


@pytest.fixture(scope="module")
def orc():
    lib = ML.load("oracle")
    if lib is None:
        pytest.fail("oracle not built: make -C oracle")
    d = lib.dll
    for fn in ("orc_tet_quality", "orc_triangle_quality", "orc_triangle_xy_quality"):
        getattr(d, fn).restype = C.c_double
        getattr(d, fn).argtypes = [C.c_void_p]
    d.orc_choose_edge_swap.restype = C.c_uint
    d.orc_choose_edge_swap.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    d.orc_swap_nint_edges.restype = C.c_uint
    d.orc_invert_map.argtypes = [C.c_uint, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    return lib


def _tet(d, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return d.orc_tet_quality(x.ctypes.data)


def _tri(d, x, xy=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return (d.orc_triangle_xy_quality if xy else d.orc_triangle_quality)(x.ctypes.data)


# --- independent numpy statement of the quality measures (quality.hpp:15-89)
def np_tet_quality(x):
    x = np.asarray(x, dtype=np.float64)

    def area(a, b, c):
        return np.linalg.norm(np.cross(b - a, c - a)) / 2

    faces = [(0, 1, 2), (1, 0, 3), (2, 1, 3), (0, 2, 3)]
    arms = math.sqrt(sum(area(*x[list(f)]) ** 2 for f in faces) / 4)
    vol = np.dot(np.cross(x[1] - x[0], x[2] - x[0]), x[3] - x[0]) / 6
    perfect = (math.sqrt(2) / 12) / (math.sqrt(math.sqrt(3) / 4) ** 3)
    return vol / arms ** 1.5 / perfect


def test_kuhn_tets_closed_form(orc):
    # every tet of the Kuhn cube split is 2^(-1/4) in this measure
    d = orc.dll
    nv, x, ne, conn = C.c_uint(), C.POINTER(C.c_double)(), C.c_uint(), C.POINTER(C.c_uint)()
    d.orc_box(3, 1, C.byref(nv), C.byref(x), C.byref(ne), C.byref(conn))
    xs = np.ctypeslib.as_array(x, shape=(nv.value * 3,)).reshape(-1, 3).copy()
    cs = np.ctypeslib.as_array(conn, shape=(ne.value * 4,)).reshape(-1, 4).copy()
    d.orc_free_buf(x)
    d.orc_free_buf(conn)
    assert cs.shape == (6, 4)
    for t in cs:
        q = _tet(d, xs[t])
        assert abs(q - 2 ** -0.25) <= 1e-15, q  # a few ulps of the closed form
        assert q == 0.8408964152537141


def test_triangle_closed_forms(orc):
    d = orc.dll
    right = [[0, 0, 0], [1, 0, 0], [1, 1, 0]]
    assert _tri(d, right) == 0.8660254037844387
    assert abs(_tri(d, right) - math.sqrt(3) / 2) <= 4e-16  # = sqrt(3)/2 to an ulp
    s = math.sqrt(3) / 2
    equi = [[0, 0, 0], [1, 0, 0], [0.5, s, 0]]
    assert abs(_tri(d, equi) - 1.0) <= 4e-16
    # the xy form is signed: reversing the orientation flips the sign only
    rev = [equi[0], equi[2], equi[1]]
    assert _tri(d, rev, xy=True) == -_tri(d, equi, xy=True)
    assert _tri(d, rev) == _tri(d, equi)
    # degenerate
    assert _tri(d, [[0, 0, 0], [1, 0, 0], [2, 0, 0]]) == 0.0


def test_tet_invariances(orc):
    d = orc.dll
    reg = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float64)
    q = _tet(d, reg)
    assert abs(abs(q) - 1.0) <= 4e-16
    # orientation: swapping two vertices flips the sign exactly
    assert _tet(d, reg[[1, 0, 2, 3]]) == -q
    rng = np.random.default_rng(7)
    for _ in range(200):
        x = rng.normal(size=(4, 3))
        q = _tet(d, x)
        # power-of-two scaling and translation by representable shifts
        assert _tet(d, x * 8.0) == q
        assert abs(_tet(d, x + 0.25) - q) <= 1e-12 * max(1.0, abs(q)) * 1e3
        # rotation invariance and the independent numpy statement
        a = rng.normal(size=(3, 3))
        qm, _ = np.linalg.qr(a)
        if np.linalg.det(qm) < 0:
            qm[:, 0] *= -1
        assert abs(_tet(d, x @ qm.T) - q) <= 1e-9 * max(1.0, abs(q))
        assert abs(np_tet_quality(x) - q) <= 1e-12 * max(1.0, abs(q)) * 1e2
    # degenerate (coplanar) tet
    assert _tet(d, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]) == 0.0


def test_invert_map_is_stable_counting_sort(orc):
    d = orc.dll
    rng = np.random.default_rng(3)
    for nout, nin in [(1, 0), (5, 17), (100, 1000), (1000, 300)]:
        a = rng.integers(0, nout, size=nin).astype(np.uint32)
        out = np.zeros(max(nin, 1), dtype=np.uint32)
        offs = np.zeros(nout + 1, dtype=np.uint32)
        d.orc_invert_map(nin, a.ctypes.data, nout, out.ctypes.data, offs.ctypes.data)
        assert (out[:nin] == np.argsort(a, kind="stable")).all()
        assert (offs == np.concatenate([[0], np.cumsum(np.bincount(a, minlength=nout))])).all()


def _mesh(orc, dim, n):
    x, c = ML.box(dim, n)
    m = ML.Mesh(orc, dim, x, c)
    m.derive_model_full()
    return m


@pytest.mark.parametrize("dim,n", [(2, 3), (3, 2)])
def test_adjacency_by_brute_force(orc, dim, n):
    m = _mesh(orc, dim, n)
    verts = {d: m.down(d, 0).reshape(m.count(d), d + 1) for d in range(1, dim + 1)}
    key = {d: {frozenset(v.tolist()): i for i, v in enumerate(verts[d])} for d in verts}
    deg = ML._DOWN_DEG
    for h in range(2, dim + 1):
        for l in range(1, h):
            dn = m.down(h, l).reshape(m.count(h), deg[h][l])
            for i in range(m.count(h)):
                # each low is the unique entity with a vertex subset of the high
                for e in dn[i]:
                    assert set(verts[l][e].tolist()) <= set(verts[h][i].tolist())
                assert len(set(dn[i].tolist())) == deg[h][l]
            # upward adjacency = stable counting sort of the uses by low
            offs, adj, dirs = m.up(l, h)
            flat = dn.ravel()
            order = np.argsort(flat, kind="stable")
            assert (adj == order // deg[h][l]).all() and (dirs == order % deg[h][l]).all()
            assert (offs == np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=m.count(l)))])).all()
    # vertex-to-entity upward adjacency likewise
    for h in range(1, dim + 1):
        offs, adj, dirs = m.up(0, h)
        flat = verts[h].ravel()
        order = np.argsort(flat, kind="stable")
        assert (adj == order // (h + 1)).all() and (dirs == order % (h + 1)).all()
    # star(l, dim) as a set: entities sharing an element, minus itself
    for l in range(0, dim):
        offs, adj = m.star(l, dim)
        dn = m.down(dim, l).reshape(m.count(dim), deg[dim][l])
        want = [set() for _ in range(m.count(l))]
        for row in dn:
            for a in row:
                want[a].update(int(b) for b in row if b != a)
        for i in range(m.count(l)):
            got = adj[offs[i]:offs[i + 1]].tolist()
            assert len(got) == len(set(got)) and set(got) == want[i], (l, i)
    del key


def _brute_indset(offs, adj, cand, good):
    NOT, IN, UNK = 0, 1, 2
    st = np.where(cand != 0, UNK, NOT).astype(np.uint32)
    for _ in range(100):
        old = st.copy()
        for i in range(len(st)):
            if old[i] != UNK:
                continue
            nb = adj[offs[i]:offs[i + 1]]
            if any(old[j] == IN for j in nb):
                st[i] = NOT
                continue
            yields = False
            for j in nb:
                if old[j] == NOT:
                    continue
                if good[j] > good[i] or (good[j] == good[i] and j < i):
                    yields = True
                    break
            if not yields:
                st[i] = IN
        if (st != UNK).all():
            return st
    raise AssertionError("no convergence")


@pytest.mark.parametrize("dim,n,ent_dim", [(2, 4, 1), (2, 4, 0), (3, 2, 1), (3, 2, 0)])
def test_indset_by_brute_force(orc, dim, n, ent_dim):
    m = _mesh(orc, dim, n)
    ne = m.count(ent_dim)
    offs, adj = m.star(ent_dim, dim)
    rng = np.random.default_rng(11 + dim + ent_dim)
    for trial in range(6):
        cand = (rng.random(ne) < 0.7).astype(np.uint32)
        # few distinct values, so equal-goodness ties are frequent
        good = rng.integers(0, 3 if trial % 2 else 50, size=ne).astype(np.float64) / 7.0
        got = m.find_indset(ent_dim, cand, good)
        want = _brute_indset(offs, adj, cand, good)
        assert (got == want).all()
        ins = np.nonzero(got == 1)[0]
        insset = set(ins.tolist())
        for i in ins:  # independent, within the candidates
            assert cand[i]
            assert not insset.intersection(adj[offs[i]:offs[i + 1]].tolist())
        for i in np.nonzero(cand)[0]:  # maximal
            if got[i] != 1:
                assert insset.intersection(adj[offs[i]:offs[i + 1]].tolist())


def _triangulations(poly):
    if len(poly) < 3:
        yield []
        return
    a, b = poly[0], poly[-1]
    for k in range(1, len(poly) - 1):
        for left in _triangulations(poly[:k + 1]):
            for right in _triangulations(poly[k:]):
                yield left + [(a, poly[k], b)] + right


def _tri_min_quality(tri, ring, e0, e1):
    # both tets of a ring triangle: (a, b, c, e1) and (b, a, c, e0)
    a, b, c = (ring[i] for i in tri)
    return min(np_tet_quality([a, b, c, e1]), np_tet_quality([b, a, c, e0]))


@pytest.mark.parametrize("rs", [3, 4, 5, 6, 7])
def test_choose_edge_swap_is_exhaustive_optimum(orc, rs):
    d = orc.dll
    k = d.orc_swap_nint_edges(rs)
    assert k == rs - 3
    catalan = {3: 1, 4: 2, 5: 5, 6: 14, 7: 42}[rs]
    tris_all = list(_triangulations(list(range(rs))))
    assert len(tris_all) == catalan
    rng = np.random.default_rng(100 + rs)
    checked = 0
    for _ in range(20):
        ang = np.sort(rng.uniform(0, 2 * np.pi, rs))
        rad = rng.uniform(0.5, 1.5, rs)
        ring = np.stack([rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-0.2, 0.2, rs)], axis=1)
        e = np.array([[0, 0, -1.0], [0, 0, 1.0]]) + rng.uniform(-0.1, 0.1, (2, 3))
        q = C.c_double()
        code = d.orc_choose_edge_swap(rs, np.ascontiguousarray(e).ctypes.data,
                                      np.ascontiguousarray(ring).ctypes.data, C.byref(q))
        scores = [min(_tri_min_quality(t, ring, e[0], e[1]) for t in tri) for tri in tris_all]
        best = max(scores)
        if best <= 0:
            continue
        assert abs(q.value - best) <= 1e-12 * abs(best)
        # the chosen code's interior edges identify a triangulation scoring best
        buf = (C.c_uint * max(2 * k, 1))()
        d.orc_swap_int_edges(rs, code, buf)
        diag = {frozenset((buf[2 * i], buf[2 * i + 1])) for i in range(k)}
        match = [s for tri, s in zip(tris_all, scores)
                 if {frozenset(p) for t in tri for p in itertools.combinations(t, 2)
                     if abs(p[0] - p[1]) not in (1, rs - 1)} == diag]
        assert len(match) == 1
        assert abs(match[0] - best) <= 1e-12 * abs(best)
        checked += 1
    assert checked >= 10
