"""ctypes front-end of the product's mesh seam (include/ohx.h).

The product library, ``omega_h_v1_amd/lib/libomega_h_amd.so``, exports the
seam with the prefix ``ohx_mesh_``: the mesh lives in HBM and every kernel
is a gfx950 HIP kernel.  ``Lib`` wraps any library that exports the same
seam under another prefix (the test suite's checkers do), and ``Mesh`` gives
one Python interface over all of them, so a parity test reads the same for
every pair.  ``product()`` loads the product and fails loudly if it has not
been built: there is no CPU path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRODUCT_LIB = os.path.join(HERE, "lib", "libomega_h_amd.so")

_u = C.c_uint
_d = C.c_double
_p = C.c_void_p
_s = C.c_char_p

# name -> (restype, argtypes); the mesh handle is always the first argument
_SIG = {
    "new_mesh": (_p, [_u, _u, _p, _u, _p]),
    "free": (None, [_p]),
    "derive_model_full": (None, [_p]),
    "set_full": (None, [_p]),
    "count": (_u, [_p, _u]),
    "get_down": (None, [_p, _u, _u, _p]),
    "up_nuses": (_u, [_p, _u, _u]),
    "get_up": (None, [_p, _u, _u, _p, _p, _p]),
    "star_nadj": (_u, [_p, _u, _u]),
    "get_star": (None, [_p, _u, _u, _p, _p]),
    "get_dual": (None, [_p, _p]),
    "get_tag": (C.c_int, [_p, _u, _s, C.c_int, _p]),
    "ntags": (_u, [_p, _u]),
    "tag_name": (_s, [_p, _u, _u]),
    "add_field": (None, [_p, _u, _s, _u, _p]),
    "free_field": (None, [_p, _u, _s]),
    "add_field_tt": (None, [_p, _u, _s, _u, _p, C.c_int]),
    "element_gradients": (None, [_p, _s]),
    "recover_by_volume": (None, [_p, _s]),
    "size_from_hessian": (None, [_p, _s, _p, _d, _d]),
    "identity_size": (None, [_p, _s]),
    "min_quality": (_d, [_p]),
    "qualities": (None, [_p, _p]),
    "measure_edges": (None, [_p, _p]),
    "find_indset": (None, [_p, _u, _p, _p, _p]),
    "refine_by_size": (_u, [_p, _d]),
    "uniformly_refine": (None, [_p]),
    "coarsen_by_size": (_u, [_p, _d, _d]),
    "coarsen_slivers": (_u, [_p, _d, _u]),
    "swap_slivers": (_u, [_p, _d, _u]),
    "adapt_timed": (C.c_int, [_p, _d, _d, _u, _u, _p]),
    "warp_to_limit": (_u, [_p, _d]),
    "metric_from_hessian": (None, [_p, _s, _d, _d, _d, _s]),
}
_DOWN_DEG = [[1, 0, 0, 0], [2, 1, 0, 0], [3, 3, 1, 0], [4, 6, 4, 1]]
# per-prefix symbol names that differ from the common name
_ALIAS = {"ohx_mesh_": {"new_mesh": "new"}}


class Lib:
    """One loaded library with a given symbol prefix."""

    def __init__(self, path, prefix):
        self.path = path
        self.prefix = prefix
        self.dll = C.CDLL(path)
        self.f = {}
        for name, (res, args) in _SIG.items():
            sym = _ALIAS.get(prefix, {}).get(name, name)
            fn = getattr(self.dll, prefix + sym, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
            self.f[name] = fn
        if prefix == "ohx_mesh_":
            self._wrap_errors()
        adapt = getattr(self.dll, prefix + "adapt", None)
        if adapt is not None:
            adapt.restype = C.c_int if prefix != "ref_" else _u
            adapt.argtypes = [_p, _d, _d, _u, _u]
            self.f["adapt"] = adapt
            if prefix == "ohx_mesh_":
                self._wrap_errors_one("adapt")

    def has(self, name):
        return name in self.f

    def _wrap_errors(self):
        """The product records errors instead of aborting; surface them."""
        last = self.dll.ohx_last_error
        last.restype = C.c_char_p
        clear = self.dll.ohx_clear_error

        def wrap(fn, name):
            def call(*a):
                clear()
                r = fn(*a)
                msg = last()
                if msg:
                    raise RuntimeError("%s: %s" % (name, msg.decode()))
                return r
            return call
        self._wrap = wrap
        for name in list(self.f):
            self.f[name] = wrap(self.f[name], name)

    def _wrap_errors_one(self, name):
        self.f[name] = self._wrap(self.f[name], name)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Mesh:
    """A mesh owned by one of the libraries."""

    def __init__(self, lib, dim, coords, conn):
        self.lib = lib
        self.dim = dim
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.uint32)
        self.h = lib.f["new_mesh"](dim, len(coords), _ptr(coords), len(conn) // 1
                                   if conn.ndim == 1 else len(conn), _ptr(conn))
        if not self.h:
            raise RuntimeError("mesh creation failed in " + lib.path)

    @classmethod
    def box_on_device(cls, lib, dim, n):
        """The product's own device-side Kuhn generator (ohx_mesh_new_box)."""
        fn = lib.dll.ohx_mesh_new_box
        fn.restype = C.c_void_p
        fn.argtypes = [_u, _u]
        self = cls.__new__(cls)
        self.lib = lib
        self.dim = dim
        self.h = fn(dim, n)
        if not self.h:
            raise RuntimeError("ohx_mesh_new_box failed: %s" % lib.dll.ohx_last_error())
        return self

    def clone(self):
        fn = self.lib.dll.ohx_mesh_clone
        fn.restype = C.c_void_p
        fn.argtypes = [_p]
        other = Mesh.__new__(Mesh)
        other.lib = self.lib
        other.dim = self.dim
        other.h = fn(self.h)
        return other

    def __del__(self):
        try:
            if self.h:
                self.lib.f["free"](self.h)
                self.h = None
        except Exception:
            pass

    def call(self, name, *args):
        return self.lib.f[name](self.h, *args)

    def derive_model_full(self):
        self.call("derive_model_full")

    def count(self, d):
        return int(self.call("count", d))

    def counts(self):
        return [self.count(d) for d in range(self.dim + 1)]

    def down(self, h, l):
        out = np.empty(self.count(h) * _DOWN_DEG[h][l], dtype=np.uint32)
        self.call("get_down", h, l, _ptr(out))
        return out

    def up(self, l, h):
        n = self.count(l)
        nuses = self.call("up_nuses", l, h)
        offs = np.empty(n + 1, dtype=np.uint32)
        adj = np.empty(max(nuses, 1), dtype=np.uint32)
        dirs = np.empty(max(nuses, 1), dtype=np.uint32)
        self.call("get_up", l, h, _ptr(offs), _ptr(adj), _ptr(dirs))
        return offs, adj[:nuses], dirs[:nuses]

    def star(self, l, h):
        n = self.count(l)
        nadj = self.call("star_nadj", l, h)
        offs = np.empty(n + 1, dtype=np.uint32)
        adj = np.empty(max(nadj, 1), dtype=np.uint32)
        self.call("get_star", l, h, _ptr(offs), _ptr(adj))
        return offs, adj[:nadj]

    def dual(self):
        d = self.dim
        out = np.empty(self.count(d) * _DOWN_DEG[d][d - 1], dtype=np.uint32)
        self.call("get_dual", _ptr(out))
        return out

    def tag(self, dim, name, f64=False):
        nc = self.call("get_tag", dim, name.encode(), int(f64), None)
        if nc < 0:
            return None
        out = np.empty(self.count(dim) * nc, dtype=np.float64 if f64 else np.uint32)
        self.call("get_tag", dim, name.encode(), int(f64), _ptr(out))
        return out

    def tag_names(self, dim):
        return [self.call("tag_name", dim, i).decode() for i in range(self.call("ntags", dim))]

    def add_field(self, dim, name, data, transfer=None):
        """A float64 field; transfer (enum osh_transfer: 1 pointwise,
        2 conserve) makes an element field follow the mesh operations."""
        data = np.ascontiguousarray(data, dtype=np.float64)
        ncomps = data.size // self.count(dim)
        if transfer is None:
            self.call("add_field", dim, name.encode(), ncomps, _ptr(data))
        else:
            self.call("add_field_tt", dim, name.encode(), ncomps, _ptr(data), int(transfer))

    def free_field(self, dim, name):
        self.call("free_field", dim, name.encode())

    def field(self, dim, name):
        return self.tag(dim, name, f64=True)

    def coords(self):
        return self.field(0, "coordinates")

    def min_quality(self):
        return float(self.call("min_quality"))

    def qualities(self):
        out = np.empty(self.count(self.dim), dtype=np.float64)
        self.call("qualities", _ptr(out))
        return out

    def measure_edges(self):
        out = np.empty(self.count(1), dtype=np.float64)
        self.call("measure_edges", _ptr(out))
        return out

    def find_indset(self, ent_dim, candidates, goodness):
        c = np.ascontiguousarray(candidates, dtype=np.uint32)
        g = np.ascontiguousarray(goodness, dtype=np.float64)
        out = np.empty(self.count(ent_dim), dtype=np.uint32)
        self.call("find_indset", ent_dim, _ptr(c), _ptr(g), _ptr(out))
        return out

    def refine_by_size(self, qual_floor):
        return int(self.call("refine_by_size", qual_floor))

    def uniformly_refine(self):
        self.call("uniformly_refine")

    def coarsen_by_size(self, qual_floor, size_floor):
        return int(self.call("coarsen_by_size", qual_floor, size_floor))

    def coarsen_slivers(self, qual_floor, nlayers):
        return int(self.call("coarsen_slivers", qual_floor, nlayers))

    def swap_slivers(self, good_qual, nlayers):
        return int(self.call("swap_slivers", good_qual, nlayers))

    def adapt(self, size_floor=1.0 / 3.0, good_qual=0.3, nlayers=4, max_ops=50):
        return int(self.call("adapt", size_floor, good_qual, nlayers, max_ops))

    def adapt_timed(self, size_floor=1.0 / 3.0, good_qual=0.3, nlayers=4, max_ops=50):
        stats = np.zeros(4, dtype=np.float64)
        r = self.call("adapt_timed", size_floor, good_qual, nlayers, max_ops, _ptr(stats))
        return int(r), stats

    # ---- many meshes at once (product only)
    @staticmethod
    def adapt_many(lib, meshes, size_floor=1.0 / 3.0, good_qual=0.3, nlayers=4, max_ops=50, nthreads=4):
        d = lib.dll
        f = d.ohx_adapt_many
        f.restype = C.c_int
        f.argtypes = [_u, C.POINTER(C.c_void_p), _d, _d, _u, _u, _u, C.c_void_p, C.c_void_p]
        arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
        res = np.zeros(len(meshes), dtype=np.int32)
        stats = np.zeros(4, dtype=np.float64)
        d.ohx_clear_error()
        r = f(len(meshes), arr, size_floor, good_qual, nlayers, max_ops, nthreads, _ptr(res), _ptr(stats))
        msg = d.ohx_last_error()
        if msg:
            raise RuntimeError("ohx_adapt_many: %s" % msg)
        return int(r), res, stats

    # ---- batches (product only)
    @classmethod
    def batch(cls, lib, meshes):
        """The product's batch of full, classified meshes (ohx_mesh_batch_new)."""
        d = lib.dll
        d.ohx_mesh_batch_new.restype = C.c_void_p
        d.ohx_mesh_batch_new.argtypes = [_u, C.POINTER(C.c_void_p)]
        arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
        self = cls.__new__(cls)
        self.lib = lib
        self.dim = meshes[0].dim
        self.h = d.ohx_mesh_batch_new(len(meshes), arr)
        if not self.h:
            raise RuntimeError("ohx_mesh_batch_new failed: %s" % d.ohx_last_error())
        return self

    def batch_size(self):
        f = self.lib.dll.ohx_mesh_batch_size
        f.restype, f.argtypes = _u, [_p]
        return int(f(self.h))

    def batch_extract(self, c):
        d = self.lib.dll
        d.ohx_mesh_batch_extract.restype = C.c_void_p
        d.ohx_mesh_batch_extract.argtypes = [_p, _u]
        other = Mesh.__new__(Mesh)
        other.lib = self.lib
        other.dim = self.dim
        other.h = d.ohx_mesh_batch_extract(self.h, c)
        if not other.h:
            raise RuntimeError("ohx_mesh_batch_extract failed: %s" % d.ohx_last_error())
        return other

    def batch_adapt(self, size_floor=1.0 / 3.0, good_qual=0.3, nlayers=4, max_ops=50):
        d = self.lib.dll
        f = d.ohx_mesh_batch_adapt
        f.restype = C.c_int
        f.argtypes = [_p, _d, _d, _u, _u, C.c_void_p, C.c_void_p]
        res = np.zeros(self.batch_size(), dtype=np.int32)
        stats = np.zeros(4, dtype=np.float64)
        d.ohx_clear_error()
        r = f(self.h, size_floor, good_qual, nlayers, max_ops, _ptr(res), _ptr(stats))
        msg = d.ohx_last_error()
        if msg:
            raise RuntimeError("ohx_mesh_batch_adapt: %s" % msg)
        return int(r), res, stats

    # ---- VTU files (product and reference)
    @classmethod
    def read_vtk(cls, lib, path):
        """osh_read_vtk's mesh: ohx_mesh_read_vtk (product) or
        read_mesh_vtk (reference)."""
        d = lib.dll
        if lib.prefix == "ohx_mesh_":
            f, dimf = d.ohx_mesh_read_vtk, d.osh_dim
        else:
            f, dimf = d.ref_read_vtk, d.ref_dim
        f.restype, f.argtypes = C.c_void_p, [_s]
        dimf.restype, dimf.argtypes = _u, [_p]
        if lib.prefix == "ohx_mesh_":
            d.ohx_clear_error()
        h = f(path.encode())
        if not h:
            msg = d.ohx_last_error() if lib.prefix == "ohx_mesh_" else b"?"
            raise RuntimeError("read_vtk %s: %s" % (path, msg))
        self = cls.__new__(cls)
        self.lib = lib
        self.h = h
        self.dim = int(dimf(h))
        return self

    def write_vtk(self, path):
        d = self.lib.dll
        if self.lib.prefix == "ohx_mesh_":
            f = d.ohx_mesh_write_vtk
            f.restype, f.argtypes = C.c_int, [_p, _s]
            d.ohx_clear_error()
            if f(self.h, path.encode()) != 0:
                raise RuntimeError("write_vtk %s: %s" % (path, d.ohx_last_error()))
        else:
            f = d.ref_write_vtk
            f.restype, f.argtypes = None, [_p, _s]
            f(self.h, path.encode())

    # ---- size-field producers
    def element_gradients(self, name):
        self.call("element_gradients", name.encode())

    def recover_by_volume(self, name):
        self.call("recover_by_volume", name.encode())

    def size_from_hessian(self, hess_name, weights, min_h, max_h):
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        self.call("size_from_hessian", hess_name.encode(), None if w is None else _ptr(w), min_h, max_h)

    def identity_size(self, name):
        self.call("identity_size", name.encode())

    def metric_from_hessian(self, hessian, eps, hmin, hmax, out="metric"):
        self.call("metric_from_hessian", hessian.encode(), eps, hmin, hmax, out.encode())

    def warp_to_limit(self, qual_floor):
        return int(self.call("warp_to_limit", qual_floor))

    def state(self):
        """Everything a topology parity check compares: connectivity of every
        dimension, every tag of every dimension."""
        st = {}
        for d in range(self.dim + 1):
            st["count%d" % d] = self.count(d)
            if d:
                st["down%d0" % d] = self.down(d, 0)
            for name in self.tag_names(d):
                a = self.tag(d, name, f64=False)
                if a is None:
                    a = self.tag(d, name, f64=True)
                st["tag%d:%s" % (d, name)] = a
        return st


def box(dim, n):
    """Kuhn box lattice: n^dim cubes (squares), each split as the
    reference's single box (tables.cpp:15-47); numpy twin of orc_box."""
    np1 = n + 1
    if dim == 3:
        corner = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                           [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
        split = np.array([[0, 1, 2, 6], [2, 3, 0, 6], [0, 3, 7, 6],
                          [7, 4, 0, 6], [0, 4, 5, 6], [5, 1, 0, 6]])
        v = np.arange(np1 ** 3)
        coords = np.stack([(v % np1) / n, ((v // np1) % np1) / n, (v // (np1 * np1)) / n], 1)
        c = np.arange(n ** 3)
        ci, cj, ck = c % n, (c // n) % n, c // (n * n)
        vid = (ci[:, None] + corner[None, :, 0]) + np1 * ((cj[:, None] + corner[None, :, 1])
                                                          + np1 * (ck[:, None] + corner[None, :, 2]))
        conn = vid[:, split].reshape(-1, 4)
    else:
        corner = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
        split = np.array([[0, 1, 2], [2, 3, 0]])
        v = np.arange(np1 ** 2)
        coords = np.stack([(v % np1) / n, (v // np1) / n, np.zeros(len(v))], 1)
        c = np.arange(n * n)
        ci, cj = c % n, c // n
        vid = (ci[:, None] + corner[None, :, 0]) + np1 * (cj[:, None] + corner[None, :, 1])
        conn = vid[:, split].reshape(-1, 3)
    return np.ascontiguousarray(coords, dtype=np.float64), np.ascontiguousarray(conn, dtype=np.uint32)


def product():
    """The product library (prefix ohx_mesh_); raises if it is not built."""
    if not os.path.exists(PRODUCT_LIB):
        raise RuntimeError("omega_h_v1_amd: %s is not built (run __graft_entry__.build())" % PRODUCT_LIB)
    return Lib(PRODUCT_LIB, "ohx_mesh_")
