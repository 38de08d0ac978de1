// The VTU file format of the public osh_read_vtk / osh_write_vtk
// (vtk_io.cpp:143-543, 571-830; base64.cpp:102-130): one XML piece per
// file, every array written "binary" -- a base64 header holding the byte
// count (or the zlib header {1, size, size, compressed size} when
// compressing) followed by the base64 data.  Host code: arrays move between
// HBM and the file through host buffers.  The layout, element order and
// attribute text follow the reference, so files are exchanged both ways
// and an uncompressed file is byte-identical to the reference's.
#include <zlib.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ohx_mesh.hpp"

namespace ohx {

namespace {

char const kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

int b64_value(int c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  if (c == '=') return 0;  // padding decodes as zero bits
  return -1;
}

// padded base64 of n bytes
std::string b64_encode(void const* data, size_t n) {
  unsigned char const* in = static_cast<unsigned char const*>(data);
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    unsigned v = (in[i] << 16) | (in[i + 1] << 8) | in[i + 2];
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (n - i == 1) {
    unsigned v = in[i] << 16;
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += "==";
  } else if (n - i == 2) {
    unsigned v = (in[i] << 16) | (in[i + 1] << 8);
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += '=';
  }
  return out;
}

// decode the groups that carry `bytes` bytes, advancing p
std::vector<unsigned char> b64_decode(char const*& p, char const* end, size_t bytes) {
  size_t groups = (bytes + 2) / 3;
  OHX_REQUIRE((size_t) (end - p) >= groups * 4, "vtu: truncated base64 data");
  std::vector<unsigned char> out(groups * 3);
  for (size_t g = 0; g < groups; ++g) {
    int v[4];
    for (int k = 0; k < 4; ++k) {
      v[k] = b64_value((unsigned char) p[k]);
      OHX_REQUIRE(v[k] >= 0, "vtu: bad base64 character");
    }
    out[3 * g + 0] = (unsigned char) ((v[0] << 2) | (v[1] >> 4));
    out[3 * g + 1] = (unsigned char) (((v[1] & 15) << 4) | (v[2] >> 2));
    out[3 * g + 2] = (unsigned char) (((v[2] & 3) << 6) | v[3]);
    p += 4;
  }
  out.resize(bytes);
  return out;
}

enum VType { V_U8, V_U32, V_U64, V_F64 };
char const* vtype_name(VType t) {
  switch (t) {
    case V_U8: return "UInt8";
    case V_U32: return "UInt32";
    case V_U64: return "UInt64";
    case V_F64: return "Float64";
  }
  return "";
}
size_t vtype_size(VType t) { return t == V_U8 ? 1 : (t == V_U32 ? 4 : 8); }

struct Writer {
  FILE* f;
  bool compress;
  void array(VType t, char const* name, u32 ncomps, void const* host, size_t nbytes) {
    std::fprintf(f, "<DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%u\" format=\"binary\">\n",
                 vtype_name(t), name, ncomps);
    if (compress) {
      uLongf clen = compressBound((uLong) nbytes);
      std::vector<unsigned char> comp(clen);
      OHX_REQUIRE(::compress2(comp.data(), &clen, static_cast<Bytef const*>(host), (uLong) nbytes,
                              Z_DEFAULT_COMPRESSION) == Z_OK,
                  "vtu: zlib compression failed");
      unsigned header[4] = {1u, (unsigned) nbytes, (unsigned) nbytes, (unsigned) clen};
      std::fputs(b64_encode(header, sizeof(header)).c_str(), f);
      std::fputs(b64_encode(comp.data(), clen).c_str(), f);
    } else {
      unsigned size = (unsigned) nbytes;
      std::fputs(b64_encode(&size, sizeof(size)).c_str(), f);
      std::fputs(b64_encode(host, nbytes).c_str(), f);
    }
    std::fputc('\n', f);
    std::fputs("</DataArray>\n", f);
  }
  void tag(Tag& t, u32 n) {
    size_t k = (size_t) n * t.ncomps;
    if (t.type == TagType::F64) {
      std::vector<double> h(k);
      to_host(h.data(), t.f.get(), k);
      array(V_F64, t.name.c_str(), t.ncomps, h.data(), k * 8);
    } else {
      std::vector<u32> h(k);
      to_host(h.data(), t.u.get(), k);
      array(V_U32, t.name.c_str(), t.ncomps, h.data(), k * 4);
    }
  }
  void connectivity(Mesh& m, u32 d) {
    size_t k = (size_t) count(m, d) * (d + 1);
    std::vector<u32> h(k);
    to_host(h.data(), ask_down(m, d, 0), k);
    array(V_U32, "connectivity", 1, h.data(), k * 4);
  }
  void tags_of(Mesh& m, u32 d, char const* except) {
    for (auto& tp : m.tags[d])
      if (!except || tp->name != except) tag(*tp, count(m, d));
  }
};

// ------------------------------------------------------------------ read
struct Reader {
  std::string text;
  size_t pos = 0;
  bool little = true, compressed = false;
  // the next line starting with prefix (leading text of the line), or false
  bool seek(char const* prefix, std::string* line) {
    size_t pl = std::strlen(prefix);
    while (pos < text.size()) {
      size_t e = text.find('\n', pos);
      if (e == std::string::npos) e = text.size();
      bool hit = text.compare(pos, pl, prefix) == 0;
      if (hit && line) *line = text.substr(pos, e - pos);
      pos = e < text.size() ? e + 1 : e;
      if (hit) return true;
    }
    return false;
  }
  // the next line, if it starts with prefix
  bool next_is(char const* prefix, std::string* line) {
    if (pos >= text.size()) return false;
    size_t e = text.find('\n', pos);
    if (e == std::string::npos) e = text.size();
    bool hit = text.compare(pos, std::strlen(prefix), prefix) == 0;
    if (line) *line = text.substr(pos, e - pos);
    pos = e < text.size() ? e + 1 : e;
    return hit;
  }
};

bool attrib(std::string const& elem, char const* name, std::string* val) {
  std::string key = std::string(name) + "=\"";
  size_t p = elem.find(key);
  if (p == std::string::npos) return false;
  p += key.size();
  size_t q = elem.find('"', p);
  OHX_REQUIRE(q != std::string::npos, "vtu: unterminated attribute");
  *val = elem.substr(p, q - p);
  return true;
}

u32 int_attrib(std::string const& elem, char const* name, bool* found = nullptr) {
  std::string v;
  bool ok = attrib(elem, name, &v);
  if (found)
    *found = ok;
  else
    OHX_REQUIRE(ok, "vtu: missing attribute");
  return ok ? (u32) std::strtoul(v.c_str(), nullptr, 10) : 0;
}

struct Array {
  VType type;
  std::string name;
  u32 ncomps;
  std::vector<unsigned char> bytes;
};

// read the array whose header line is `line`, holding n tuples
Array read_array(Reader& r, std::string const& line, u32 n) {
  Array a;
  std::string v;
  OHX_REQUIRE(attrib(line, "type", &v), "vtu: array without a type");
  if (v == "UInt8")
    a.type = V_U8;
  else if (v == "UInt32")
    a.type = V_U32;
  else if (v == "UInt64")
    a.type = V_U64;
  else if (v == "Float64")
    a.type = V_F64;
  else
    fail("vtu: unsupported array type", __FILE__, __LINE__);
  OHX_REQUIRE(attrib(line, "Name", &a.name), "vtu: array without a name");
  bool has_nc = false;
  a.ncomps = int_attrib(line, "NumberOfComponents", &has_nc);
  if (!has_nc) a.ncomps = 1;
  OHX_REQUIRE(attrib(line, "format", &v), "vtu: array without a format");
  size_t nbytes = (size_t) n * a.ncomps * vtype_size(a.type);
  if (v == "binary") {
    // the base64 text runs up to the first character outside the alphabet
    size_t s = r.pos;
    while (s < r.text.size() && std::isspace((unsigned char) r.text[s])) ++s;
    size_t e = s;
    while (e < r.text.size() && b64_value((unsigned char) r.text[e]) >= 0) ++e;
    char const* p = r.text.data() + s;
    char const* end = r.text.data() + e;
    if (r.compressed) {
      std::vector<unsigned char> h = b64_decode(p, end, 16);
      unsigned header[4];
      std::memcpy(header, h.data(), 16);
      if (!r.little)
        for (unsigned& x : header) x = __builtin_bswap32(x);
      OHX_REQUIRE(header[0] == 1u, "vtu: multi-block compressed arrays are not supported");
      std::vector<unsigned char> comp = b64_decode(p, end, header[3]);
      a.bytes.resize(nbytes);
      uLongf out = (uLongf) nbytes;
      OHX_REQUIRE(::uncompress(a.bytes.data(), &out, comp.data(), (uLong) comp.size()) == Z_OK && out == nbytes,
                  "vtu: zlib decompression failed");
    } else {
      b64_decode(p, end, 4);
      a.bytes = b64_decode(p, end, nbytes);
    }
    r.pos = e;
    if (!r.little) {
      size_t w = vtype_size(a.type);
      for (size_t i = 0; i + w <= a.bytes.size(); i += w)
        for (size_t k = 0; k < w / 2; ++k) std::swap(a.bytes[i + k], a.bytes[i + w - 1 - k]);
    }
  } else if (v == "ascii") {
    a.bytes.resize(nbytes);
    char const* p = r.text.c_str() + r.pos;
    char* q = nullptr;
    size_t k = (size_t) n * a.ncomps;
    for (size_t i = 0; i < k; ++i) {
      switch (a.type) {
        case V_U8: a.bytes[i] = (unsigned char) std::strtoul(p, &q, 10); break;
        case V_U32: {
          u32 x = (u32) std::strtoul(p, &q, 10);
          std::memcpy(&a.bytes[4 * i], &x, 4);
          break;
        }
        case V_U64: {
          unsigned long long x = std::strtoull(p, &q, 10);
          std::memcpy(&a.bytes[8 * i], &x, 8);
          break;
        }
        case V_F64: {
          double x = std::strtod(p, &q);
          std::memcpy(&a.bytes[8 * i], &x, 8);
          break;
        }
      }
      OHX_REQUIRE(q != p, "vtu: short ascii array");
      p = q;
    }
    r.pos = p - r.text.c_str();
  } else {
    fail("vtu: unknown array format", __FILE__, __LINE__);
  }
  OHX_REQUIRE(r.seek("</DataArray", nullptr), "vtu: array not closed");
  return a;
}

// as a tag of dimension d: Float64 stays, the integer types become UInt32
void add_array_tag(Mesh& m, u32 d, Array const& a, u32 n) {
  size_t k = (size_t) n * a.ncomps;
  if (a.type == V_F64) {
    add_tag_f64(m, d, a.name.c_str(), a.ncomps, to_device(reinterpret_cast<double const*>(a.bytes.data()), k));
    return;
  }
  std::vector<u32> h(k);
  for (size_t i = 0; i < k; ++i) {
    unsigned long long x = 0;
    std::memcpy(&x, &a.bytes[i * vtype_size(a.type)], vtype_size(a.type));
    OHX_REQUIRE(x <= 0xffffffffull, "vtu: integer array value beyond 32 bits");
    h[i] = (u32) x;
  }
  add_tag_u32(m, d, a.name.c_str(), a.ncomps, to_device(h.data(), k));
}

// the tags of the section that starts with prefix
void read_tags(Reader& r, Mesh& m, u32 d, char const* prefix) {
  OHX_REQUIRE(r.seek(prefix, nullptr), "vtu: missing data section");
  std::string line;
  u32 n = count(m, d);
  while (r.next_is("<DataArray", &line)) add_array_tag(m, d, read_array(r, line, n), n);
}

void read_conn(Reader& r, Mesh& m, u32 d, u32 n, char const* section) {
  OHX_REQUIRE(r.seek(section, nullptr), "vtu: missing connectivity section");
  std::string line;
  OHX_REQUIRE(r.seek("<DataArray", &line), "vtu: missing connectivity");
  Array a = read_array(r, line, n * (d + 1));
  OHX_REQUIRE(a.type == V_U32 && a.name == "connectivity" && a.ncomps == 1, "vtu: bad connectivity array");
  set_ents(m, d, n, to_device(reinterpret_cast<u32 const*>(a.bytes.data()), (size_t) n * (d + 1)));
}

std::string vtu_path(char const* path, bool* parallel) {
  std::string p(path);
  size_t dot = p.rfind('.');
  size_t slash = p.rfind('/');
  OHX_REQUIRE(dot != std::string::npos && (slash == std::string::npos || dot > slash),
              "vtk: the file name needs a .vtu or .pvtu suffix");
  std::string suffix = p.substr(dot + 1);
  *parallel = suffix == "pvtu";
  // one piece: the piece file is the prefix with ".vtu" (files.cpp:45-72)
  return p.substr(0, dot) + ".vtu";
}

}  // namespace

bool g_vtk_compress = false;

// write_vtu_opts (vtk_io.cpp:496-543) and write_pvtu (741-774)
void write_vtk(Mesh& m, char const* path) {
  bool parallel = false;
  std::string piece = vtu_path(path, &parallel);
  u32 dim = m.dim;
  bool edges = dim > 1 && has_dim(m, 1), faces = dim > 2 && has_dim(m, 2);
  FILE* f = std::fopen(piece.c_str(), "w");
  OHX_REQUIRE(f, "vtk: cannot open the output file");
  Writer w{f, g_vtk_compress};
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" byte_order=\"LittleEndian\" header_type=\"UInt32\"%s>\n",
               g_vtk_compress ? " compressor=\"vtkZLibDataCompressor\"" : "");
  std::fprintf(f, "<UnstructuredGrid>\n");
  std::fprintf(f, "<Piece NumberOfPoints=\"%u\" NumberOfCells=\"%u\"", count(m, 0), count(m, dim));
  if (edges) std::fprintf(f, " NumberOfEdges=\"%u\"", count(m, 1));
  if (faces) std::fprintf(f, " NumberOfFaces=\"%u\"", count(m, 2));
  std::fprintf(f, " Rep=\"%s\">\n", m.full ? "Full" : "Reduced");
  std::fprintf(f, "<Points>\n");
  Tag* xt = find_tag(m, 0, "coordinates");
  OHX_REQUIRE(xt, "vtk: the mesh has no coordinates");
  w.tag(*xt, count(m, 0));
  std::fprintf(f, "</Points>\n");
  if (edges) {
    std::fprintf(f, "<Edges>\n");
    w.connectivity(m, 1);
    std::fprintf(f, "</Edges>\n");
  }
  if (faces) {
    std::fprintf(f, "<Faces>\n");
    w.connectivity(m, 2);
    std::fprintf(f, "</Faces>\n");
  }
  std::fprintf(f, "<Cells>\n");
  w.connectivity(m, dim);
  u32 ne = count(m, dim), vpe = dim + 1;
  std::vector<u32> offs(ne);
  for (u32 i = 0; i < ne; ++i) offs[i] = (i + 1) * vpe;
  w.array(V_U32, "offsets", 1, offs.data(), (size_t) ne * 4);
  static unsigned char const types[4] = {1, 3, 5, 10};  // VTK vertex, line, triangle, tetra
  std::vector<unsigned char> ty(ne, types[dim]);
  w.array(V_U8, "types", 1, ty.data(), ne);
  std::fprintf(f, "</Cells>\n");
  std::fprintf(f, "<PointData>\n");
  w.tags_of(m, 0, "coordinates");
  std::fprintf(f, "</PointData>\n");
  if (edges) {
    std::fprintf(f, "<EdgeData>\n");
    w.tags_of(m, 1, nullptr);
    std::fprintf(f, "</EdgeData>\n");
  }
  if (faces) {
    std::fprintf(f, "<FaceData>\n");
    w.tags_of(m, 2, nullptr);
    std::fprintf(f, "</FaceData>\n");
  }
  std::fprintf(f, "<CellData>\n");
  w.tags_of(m, dim, nullptr);
  std::fprintf(f, "</CellData>\n");
  std::fprintf(f, "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  std::fclose(f);
  if (!parallel) return;
  FILE* pf = std::fopen(path, "w");
  OHX_REQUIRE(pf, "vtk: cannot open the output file");
  auto describe = [&](Tag& t) {
    std::fprintf(pf, "<PDataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%u\" format=\"ascii\"/>\n",
                 t.type == TagType::F64 ? "Float64" : "UInt32", t.name.c_str(), t.ncomps);
  };
  std::fprintf(pf, "<VTKFile type=\"PUnstructuredGrid\">\n<PUnstructuredGrid GhostLevel=\"0\">\n<PPointData>\n");
  for (auto& tp : m.tags[0])
    if (tp->name != "coordinates") describe(*tp);
  std::fprintf(pf, "</PPointData>\n<PCellData>\n");
  for (auto& tp : m.tags[dim]) describe(*tp);
  std::fprintf(pf, "</PCellData>\n<PPoints>\n");
  describe(*xt);
  std::string base = piece.substr(piece.rfind('/') == std::string::npos ? 0 : piece.rfind('/') + 1);
  std::fprintf(pf, "</PPoints>\n<Piece Source=\"%s\"/>\n</PUnstructuredGrid>\n</VTKFile>\n", base.c_str());
  std::fclose(pf);
}

// read_vtu_opts (vtk_io.cpp:671-727)
void read_vtk(char const* path, Mesh& m) {
  bool parallel = false;
  std::string piece = vtu_path(path, &parallel);
  FILE* f = std::fopen(piece.c_str(), "rb");
  OHX_REQUIRE(f, "vtk: cannot open the input file");
  Reader r;
  std::fseek(f, 0, SEEK_END);
  long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  r.text.resize(len > 0 ? (size_t) len : 0);
  if (len > 0) OHX_REQUIRE(std::fread(&r.text[0], 1, (size_t) len, f) == (size_t) len, "vtk: short read");
  std::fclose(f);
  std::string line, v;
  OHX_REQUIRE(r.seek("<VTKFile", &line), "vtu: not a VTK file");
  r.little = !(attrib(line, "byte_order", &v) && v == "BigEndian");
  r.compressed = attrib(line, "compressor", &v);
  OHX_REQUIRE(r.seek("<Piece", &line), "vtu: no piece");
  u32 nverts = int_attrib(line, "NumberOfPoints");
  u32 nelems = int_attrib(line, "NumberOfCells");
  bool do_edges = false, do_faces = false;
  u32 nedges = int_attrib(line, "NumberOfEdges", &do_edges);
  u32 nfaces = int_attrib(line, "NumberOfFaces", &do_faces);
  bool full = attrib(line, "Rep", &v) && v == "Full";
  OHX_REQUIRE(nelems > 0, "vtu: a mesh needs elements");
  // the element dimension from the cell types (read_dimension, 582-607)
  size_t after_piece = r.pos;
  OHX_REQUIRE(r.seek("<Cells", nullptr), "vtu: no cells");
  u32 dim = 4;
  while (r.seek("<DataArray", &line)) {
    std::string name;
    OHX_REQUIRE(attrib(line, "Name", &name), "vtu: array without a name");
    if (name != "types") continue;
    Array a = read_array(r, line, nelems);
    OHX_REQUIRE(a.type == V_U8 && a.ncomps == 1, "vtu: bad cell types");
    static unsigned char const types[4] = {1, 3, 5, 10};
    for (u32 d = 0; d < 4; ++d)
      if (a.bytes[0] == types[d]) dim = d;
    OHX_REQUIRE(dim <= 3, "vtu: cells are not simplices");
    for (unsigned char t : a.bytes) OHX_REQUIRE(t == types[dim], "vtu: mixed cell types");
    break;
  }
  OHX_REQUIRE(dim <= 3, "vtu: no cell types");
  m.dim = dim;
  m.full = full;
  r.pos = after_piece;
  set_ents(m, 0, nverts, Buf<u32>());
  OHX_REQUIRE(r.seek("<Points", nullptr), "vtu: no points");
  OHX_REQUIRE(r.next_is("<DataArray", &line), "vtu: no point coordinates");
  add_array_tag(m, 0, read_array(r, line, nverts), nverts);
  if (do_edges) read_conn(r, m, 1, nedges, "<Edges");
  if (do_faces) read_conn(r, m, 2, nfaces, "<Faces");
  read_conn(r, m, dim, nelems, "<Cells");
  read_tags(r, m, 0, "<PointData");
  if (dim > 1 && has_dim(m, 1) && do_edges) read_tags(r, m, 1, "<EdgeData");
  if (dim > 2 && has_dim(m, 2) && do_faces) read_tags(r, m, 2, "<FaceData");
  read_tags(r, m, dim, "<CellData");
}

}  // namespace ohx
