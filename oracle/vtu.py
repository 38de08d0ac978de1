"""CPU oracle of the VTU file format of osh_read_vtk / osh_write_vtk, and of
mesh_diff.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the product's
VTU writer and reader; never by the product.

A from-scratch Python restatement of the reference's single-piece .vtu
layout and its .pvtu index:

* file and piece headers, section order and array order --
  vtk_io.cpp:431-443 (VTKFile line), 461-474 (Piece line), 496-543
  (write_vtu_opts), 400-418 (connectivity, offsets, types);
* one array: the ``<DataArray type= Name= NumberOfComponents= format=>`` line
  (vtk_io.cpp:313-342), then the base64 of a header -- the byte count as one
  UInt32, or {1, size, size, compressed size} when compressing -- followed
  directly by the base64 of the (zlib ``compress`` default level) data, on
  one line (vtk_io.cpp:150-175, base64.cpp:102-120: standard padded base64,
  each of the two parts padded on its own);
* the .pvtu index (vtk_io.cpp:729-774) with every array described ``ascii``;
* the reader (vtk_io.cpp:176-212, 344-392, 582-727): the element dimension
  from the first cell type, tags in file order, Rep="Full" or reduced.

Pinned against files written by the compiled reference (tests/golden/vtu/):
byte-identical writes and equal arrays on reads.  mesh_diff (mesh_diff.cpp:
16-178) is pinned against the compiled reference's own mesh_diff on planted
differences: the same verdict and the same messages.
"""
import base64
import re
import zlib

import numpy as np

VTK_SIMPLEX = (1, 3, 5, 10)  # vertex, line, triangle, tetrahedron
_NP = {"UInt8": np.uint8, "UInt32": np.uint32, "UInt64": np.uint64, "Float64": np.float64}
_NAME = {np.dtype(np.uint8): "UInt8", np.dtype(np.uint32): "UInt32",
         np.dtype(np.uint64): "UInt64", np.dtype(np.float64): "Float64"}


class Piece:
    """A mesh as the file holds it: dim, rep, vertex count, connectivity of
    the written dimensions, tags per dimension in order (name, ncomps,
    array)."""

    def __init__(self, dim, full, nverts, conn, tags):
        self.dim = dim
        self.full = full
        self.nverts = nverts
        self.conn = conn        # {d: u32 array of count*(d+1)}
        self.tags = tags        # {d: [(name, ncomps, array)]}, coordinates first at d=0

    def count(self, d):
        if d == 0:
            return self.nverts
        return len(self.conn[d]) // (d + 1)


def _array(name, ncomps, a, compress):
    a = np.ascontiguousarray(a)
    raw = a.tobytes()
    head = '<DataArray type="%s" Name="%s" NumberOfComponents="%u" format="binary">\n' % (
        _NAME[a.dtype], name, ncomps)
    if compress:
        comp = zlib.compress(raw)
        hdr = np.array([1, len(raw), len(raw), len(comp)], dtype="<u4").tobytes()
        body = base64.b64encode(hdr) + base64.b64encode(comp)
    else:
        hdr = np.array([len(raw)], dtype="<u4").tobytes()
        body = base64.b64encode(hdr) + base64.b64encode(raw)
    return head.encode() + body + b"\n</DataArray>\n"


def write_vtu(p, compress=False):
    """The bytes of the .vtu piece (vtk_io.cpp:496-543)."""
    dim = p.dim
    edges = dim > 1 and 1 in p.conn
    faces = dim > 2 and 2 in p.conn
    out = [b'<VTKFile type="UnstructuredGrid" byte_order="LittleEndian" header_type="UInt32"'
           + (b' compressor="vtkZLibDataCompressor"' if compress else b"") + b">\n",
           b"<UnstructuredGrid>\n"]
    line = '<Piece NumberOfPoints="%u" NumberOfCells="%u"' % (p.nverts, p.count(dim))
    if edges:
        line += ' NumberOfEdges="%u"' % p.count(1)
    if faces:
        line += ' NumberOfFaces="%u"' % p.count(2)
    line += ' Rep="%s">\n' % ("Full" if p.full else "Reduced")
    out.append(line.encode())
    coords = [t for t in p.tags[0] if t[0] == "coordinates"][0]
    out.append(b"<Points>\n" + _array(*coords, compress) + b"</Points>\n")

    def conn(d):
        return _array("connectivity", 1, p.conn[d].astype(np.uint32), compress)
    if edges:
        out.append(b"<Edges>\n" + conn(1) + b"</Edges>\n")
    if faces:
        out.append(b"<Faces>\n" + conn(2) + b"</Faces>\n")
    ne = p.count(dim)
    offs = (np.arange(1, ne + 1, dtype=np.uint64) * (dim + 1)).astype(np.uint32)
    types = np.full(ne, VTK_SIMPLEX[dim], dtype=np.uint8)
    out.append(b"<Cells>\n" + conn(dim) + _array("offsets", 1, offs, compress)
               + _array("types", 1, types, compress) + b"</Cells>\n")

    def section(tag, d, skip):
        body = b"".join(_array(n, c, a, compress) for (n, c, a) in p.tags.get(d, []) if n != skip)
        return b"<" + tag + b">\n" + body + b"</" + tag + b">\n"
    out.append(section(b"PointData", 0, "coordinates"))
    if edges:
        out.append(section(b"EdgeData", 1, None))
    if faces:
        out.append(section(b"FaceData", 2, None))
    out.append(section(b"CellData", dim, None))
    out.append(b"</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    return b"".join(out)


def write_pvtu(p, piece_name):
    """The bytes of the .pvtu index of one piece named piece_name
    (vtk_io.cpp:729-774); GhostLevel is 0 for a serial mesh."""
    def desc(n, c, a):
        return '<PDataArray type="%s" Name="%s" NumberOfComponents="%u" format="ascii"/>\n' % (
            _NAME[np.asarray(a).dtype], n, c)
    s = '<VTKFile type="PUnstructuredGrid">\n<PUnstructuredGrid GhostLevel="0">\n<PPointData>\n'
    s += "".join(desc(*t) for t in p.tags[0] if t[0] != "coordinates")
    s += "</PPointData>\n<PCellData>\n"
    s += "".join(desc(*t) for t in p.tags.get(p.dim, []))
    s += "</PCellData>\n<PPoints>\n"
    s += desc(*[t for t in p.tags[0] if t[0] == "coordinates"][0])
    s += '</PPoints>\n<Piece Source="%s"/>\n</PUnstructuredGrid>\n</VTKFile>\n' % piece_name
    return s.encode()


def _attr(line, name):
    m = re.search(r'\b%s="([^"]*)"' % name, line)
    return m.group(1) if m else None


class _Lines:
    def __init__(self, text):
        self.lines = text.split("\n")
        self.i = 0

    def seek(self, prefix):
        while self.i < len(self.lines):
            ln = self.lines[self.i]
            self.i += 1
            if ln.startswith(prefix):
                return ln
        raise ValueError("vtu: no line starting with %s" % prefix)

    def next_is(self, prefix):
        if self.i >= len(self.lines):
            return None
        ln = self.lines[self.i]
        self.i += 1
        return ln if ln.startswith(prefix) else None


def _read_array(r, line, n, compressed):
    dt = np.dtype(_NP[_attr(line, "type")]).newbyteorder("<")
    name = _attr(line, "Name")
    nc = int(_attr(line, "NumberOfComponents") or 1)
    fmt = _attr(line, "format")
    nbytes = n * nc * dt.itemsize
    if fmt == "binary":
        text = r.lines[r.i].strip()
        r.i += 1
        if compressed:
            hdr = np.frombuffer(base64.b64decode(text[:24]), dtype="<u4")
            clen = int(hdr[3])
            comp = base64.b64decode(text[24:24 + 4 * ((clen + 2) // 3)])
            raw = zlib.decompress(comp)
        else:
            raw = base64.b64decode(text[8:8 + 4 * ((nbytes + 2) // 3)])
        a = np.frombuffer(raw[:nbytes], dtype=dt).astype(dt.newbyteorder("="))
    else:
        vals = []
        while len(vals) < n * nc:
            vals += r.lines[r.i].split()
            r.i += 1
        a = np.array(vals[:n * nc], dtype=np.float64 if dt.kind == "f" else np.uint64).astype(
            dt.newbyteorder("="))
    r.seek("</DataArray")
    return name, nc, a


def read_vtu(data):
    """A Piece from .vtu bytes (vtk_io.cpp:671-727)."""
    text = data.decode("ascii")
    r = _Lines(text)
    head = r.seek("<VTKFile")
    compressed = _attr(head, "compressor") is not None
    if _attr(head, "byte_order") == "BigEndian":
        raise ValueError("big-endian files are not covered by this oracle")
    piece = r.seek("<Piece")
    nverts = int(_attr(piece, "NumberOfPoints"))
    nelems = int(_attr(piece, "NumberOfCells"))
    nedges = _attr(piece, "NumberOfEdges")
    nfaces = _attr(piece, "NumberOfFaces")
    full = _attr(piece, "Rep") == "Full"
    start = r.i
    r.seek("<Cells")
    dim = None
    while dim is None:
        line = r.seek("<DataArray")
        if _attr(line, "Name") == "types":
            _, _, t = _read_array(r, line, nelems, compressed)
            dim = VTK_SIMPLEX.index(int(t[0]))
    r.i = start
    r.seek("<Points")
    tags = {0: [_read_array(r, r.next_is("<DataArray"), nverts, compressed)]}
    conn = {}
    sides = [(1, nedges, "<Edges", "<EdgeData")] if nedges is not None and dim > 1 else []
    sides += [(2, nfaces, "<Faces", "<FaceData")] if nfaces is not None and dim > 2 else []
    for d, n, sec, _ in sides:
        r.seek(sec)
        conn[d] = _read_array(r, r.seek("<DataArray"), int(n) * (d + 1), compressed)[2]
    r.seek("<Cells")
    conn[dim] = _read_array(r, r.seek("<DataArray"), nelems * (dim + 1), compressed)[2]
    p = Piece(dim, full, nverts, conn, tags)
    for d, sec in [(0, "<PointData")] + [(d, s) for d, _, _, s in sides] + [(dim, "<CellData")]:
        r.seek(sec)
        n = p.count(d)
        while True:
            line = r.next_is("<DataArray")
            if line is None:
                break
            tags.setdefault(d, []).append(_read_array(r, line, n, compressed))
    return p


def piece_of(mesh, full=True):
    """The Piece of a mesh of tests/meshlib.py (any library), tags in the
    mesh's own order."""
    dim = mesh.dim
    conn = {d: mesh.down(d, 0).ravel() for d in range(1, dim + 1)}
    tags = {}
    for d in range(dim + 1):
        lst = []
        n = mesh.count(d)
        for name in mesh.tag_names(d):
            a = mesh.tag(d, name, f64=False)
            if a is None:
                a = mesh.tag(d, name, f64=True)
            a = np.ascontiguousarray(a).ravel()
            lst.append((name, len(a) // n if n else 0, a))
        tags[d] = lst
    return Piece(dim, full, mesh.count(0), conn, tags)


def mesh_diff(a, b, tol=1e-6, floor=0.0, superset=False):
    """(0 or 1, the lines the reference prints) for two Pieces
    (mesh_diff.cpp:141-178).  FP64 arrays: relative difference against the
    larger magnitude unless both are below floor; the maximum is the serial
    build's, over entries 1.. only and blind to NaN (doubles.cpp:100-109)."""
    out = []
    if a.dim != b.dim:
        return 1, ["mesh dimensionalities %u and %u differ" % (a.dim, b.dim)]

    def has(p, d):
        return d == 0 or d in p.conn

    def f64_differ(x, y, n):
        with np.errstate(invalid="ignore", divide="ignore"):
            fa, fb = np.abs(x), np.abs(y)
            fm = np.where(fb > fa, fb, fa)
            diffs = np.abs(x - y) / fm
            diffs = np.where((fa < floor) & (fb < floor), 0.0, diffs)
        rest = diffs[1:n]
        rest = rest[~np.isnan(rest)]
        maxdiff = -np.finfo(np.float64).max
        if len(rest):
            maxdiff = max(maxdiff, float(rest.max()))
        if not maxdiff > tol:
            return False
        out.append("max relative difference %e" % maxdiff)
        for i in np.nonzero(diffs == maxdiff)[0]:
            out.append("max difference pair is (%e, %e), #%u" % (x[i], y[i], i))
        return True

    def tag_differ(ta, tb, n):
        (na, ca, xa), (_, cb, xb) = ta, tb
        if ca != cb:
            out.append('tag "%s" sizes %u and %u differ' % (na, ca, cb))
            return True
        if xa.dtype != xb.dtype:
            out.append('tag "%s" types differ' % na)
            return True
        if xa.dtype == np.float64:
            bad = f64_differ(xa, xb, n * ca)
        else:
            bad = not np.array_equal(xa, xb)
        if bad:
            out.append('tag "%s" contents differ' % na)
        return bad

    def tags_differ(d, n):
        tb_by = {t[0]: t for t in b.tags.get(d, [])}
        names_a = {t[0] for t in a.tags.get(d, [])}
        for ta in a.tags.get(d, []):
            tb = tb_by.get(ta[0])
            if tb is None:
                if superset:
                    continue
                out.append('tag "%s" only on first mesh' % ta[0])
                return True
            if tag_differ(ta, tb, n):
                return True
        if not superset:
            for tb in b.tags.get(d, []):
                if tb[0] not in names_a:
                    out.append('tag "%s" only on second mesh' % tb[0])
                    return True
        return False

    for d in range(a.dim + 1):
        ha, hb = has(a, d), has(b, d)
        if not ha and not hb:
            continue
        if ha and not hb:
            return 1, out + ["only first mesh has dimension %u" % d]
        if hb and not ha:
            if superset:
                continue
            return 1, out + ["only second mesh has dimension %u" % d]
        if a.count(d) != b.count(d):
            return 1, out + ["counts %u and %u differ for dimension %u" % (a.count(d), b.count(d), d)]
        n = a.count(d)
        if d and not np.array_equal(a.conn[d], b.conn[d]):
            return 1, out + ["connectivity differs for dimension %u" % d]
        if tags_differ(d, n):
            return 1, out + ["tags for dimension %u differ" % d]
    return 0, out
