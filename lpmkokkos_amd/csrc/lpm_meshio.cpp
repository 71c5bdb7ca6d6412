// Mesh and field files; see lpm_meshio.hpp.
#include "lpm_meshio.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>

namespace lpm {

void add_mesh_vars(nc::Writer& w, const TreeMesh& m, const double* phys_v, const double* phys_f, const double* area) {
  const int ndim = w.dim("ndim", 3);
  const int nverts = w.dim("nverts", m.nverts());
  const int nedges = w.dim("nedges", m.n_edges);
  const int nfaces = w.dim("nfaces", m.nfaces());
  const int nfaceverts = w.dim("nfaceverts", m.nfaceverts);
  const int two = w.dim("two", 2);
  const int four = w.dim("four", 4);
  // vertices
  w.var("phys_crds_verts", nc::kDouble, {nverts, ndim}, phys_v ? phys_v : m.vert_phys.data()->v);
  w.var("lag_crds_verts", nc::kDouble, {nverts, ndim}, m.vert_lag.data()->v);
  // edges
  w.var("edge_origs", nc::kInt, {nedges}, m.edge_orig.data());
  w.var("edge_dests", nc::kInt, {nedges}, m.edge_dest.data());
  w.var("edge_lefts", nc::kInt, {nedges}, m.edge_left.data());
  w.var("edge_rights", nc::kInt, {nedges}, m.edge_right.data());
  w.var("edge_parents", nc::kInt, {nedges}, m.edge_parent.data());
  w.var("edge_kids", nc::kInt, {nedges, two}, m.edge_kids.data());
  // faces; a face's centre particle is the face of the same index
  // (Faces::initFromSeed and Faces::divide, src/LpmFaces.cpp:51, :20)
  std::vector<int32_t> centers((size_t)m.nfaces());
  std::iota(centers.begin(), centers.end(), 0);
  w.var("phys_crds_faces", nc::kDouble, {nfaces, ndim}, phys_f ? phys_f : m.face_phys.data()->v);
  w.var("lag_crds_faces", nc::kDouble, {nfaces, ndim}, m.face_lag.data()->v);
  w.var("face_area", nc::kDouble, {nfaces}, area ? area : m.face_area.data());
  w.var("face_mask", nc::kByte, {nfaces}, m.face_mask.data());
  w.var("face_edges", nc::kInt, {nfaces, nfaceverts}, m.face_edges.data());
  w.var("face_verts", nc::kInt, {nfaces, nfaceverts}, m.face_verts.data());
  w.var_copy("face_centers", nc::kInt, {nfaces}, centers.data());
  w.var("face_tree_level", nc::kInt, {nfaces}, m.face_level.data());
  w.var("face_parents", nc::kInt, {nfaces}, m.face_parent.data());
  w.var("face_kids", nc::kInt, {nfaces, four}, m.face_kids.data());
  w.att_text("MeshSeed", m.seed_id());
  w.att_text("FaceKind", m.face_kind());
  w.att_text("Geom", "SphereGeometry");
  w.att_int("baseTreeDepth", m.level);
}

bool write_bve_nc(const std::string& path, const BveFields& s, const TreeMesh* mesh, std::string& err) {
  if (mesh && (mesh->nverts() != s.nv || mesh->nfaces() != s.nf)) {
    err = "mesh sizes differ from the state's";
    return false;
  }
  // the tree's topology describes the state only if the state's divided
  // faces are the tree's; areas are the state's own (a caller may have set
  // areas other than the tree's, and a restart must read them back)
  if (mesh && !std::equal(s.fmask.begin(), s.fmask.end(), mesh->face_mask.begin())) {
    err = "the state's face mask differs from the tree's: write it without a tree level";
    return false;
  }
  nc::Writer w;
  if (mesh) {
    add_mesh_vars(w, *mesh, s.vx.data(), s.fx.data(), s.farea.data());
  } else {
    const int ndim = w.dim("ndim", 3);
    const int nverts = w.dim("nverts", s.nv);
    const int nfaces = w.dim("nfaces", s.nf);
    w.var("phys_crds_verts", nc::kDouble, {nverts, ndim}, s.vx.data());
    w.var("phys_crds_faces", nc::kDouble, {nfaces, ndim}, s.fx.data());
    w.var("face_area", nc::kDouble, {nfaces}, s.farea.data());
    w.var("face_mask", nc::kByte, {nfaces}, s.fmask.data());
  }
  // the dimension ids: ndim is always 0, nverts 1; nfaces follows nedges
  const int ndim = 0, nverts = 1, nfaces = mesh ? 3 : 2;
  // fields, as writeScalarField / writeVectorField write them (a "units"
  // attribute on each, empty as the reference's default)
  w.var("relVortVerts", nc::kDouble, {nverts}, s.vzeta.data(), "");
  w.var("absVortVerts", nc::kDouble, {nverts}, s.vabsv.data(), "");
  w.var("streamFnVerts", nc::kDouble, {nverts}, s.vpsi.data(), "");
  w.var("velocityVerts", nc::kDouble, {nverts, ndim}, s.vu.data(), "");
  w.var("relVortFaces", nc::kDouble, {nfaces}, s.fzeta.data(), "");
  w.var("absVortFaces", nc::kDouble, {nfaces}, s.fabsv.data(), "");
  w.var("streamFnFaces", nc::kDouble, {nfaces}, s.fpsi.data(), "");
  w.var("velocityFaces", nc::kDouble, {nfaces, ndim}, s.fu.data(), "");
  // resume attributes
  w.att_double("time", s.time);
  w.att_double("Omega", s.omega);
  w.att_int("velocity_prec", s.vel_prec);
  w.att_int("velocity_form", s.vel_form);
  w.att_int("streamfn_valid", s.psi_valid ? 1 : 0);
  w.att_int("absvort_valid", s.absv_valid ? 1 : 0);
  w.att_double("absvort_omega", s.absv_omega);
  return w.write(path, err);
}

namespace {

bool att_int(const nc::Reader& r, const char* name, int& out) {
  const nc::Att* a = r.gatt(name);
  if (!a || a->type != nc::kInt || a->ints.empty()) return false;
  out = a->ints[0];
  return true;
}

bool att_double(const nc::Reader& r, const char* name, double& out) {
  const nc::Att* a = r.gatt(name);
  if (!a || a->type != nc::kDouble || a->doubles.empty()) return false;
  out = a->doubles[0];
  return true;
}

}  // namespace

bool read_bve_nc(const std::string& path, BveFields& s, std::string& err) {
  nc::Reader r;
  if (!r.open(path, err)) return false;
  s = BveFields();
  s.nv = r.dim("nverts");
  s.nf = r.dim("nfaces");
  if (s.nv < 0 || s.nf < 1 || r.dim("ndim") != 3) {
    err = path + ": missing or bad nverts / nfaces / ndim dimensions";
    return false;
  }
  auto need = [&](const char* name, std::vector<double>& out, int64_t n) {
    if (!r.read(name, out) || (int64_t)out.size() != n) {
      err = path + ": missing or mis-sized variable " + name;
      return false;
    }
    return true;
  };
  if (!need("phys_crds_verts", s.vx, 3 * s.nv) || !need("phys_crds_faces", s.fx, 3 * s.nf) ||
      !need("face_area", s.farea, s.nf) || !need("relVortVerts", s.vzeta, s.nv) || !need("relVortFaces", s.fzeta, s.nf))
    return false;
  if (!r.read("face_mask", s.fmask) || (int64_t)s.fmask.size() != s.nf) {
    err = path + ": missing or mis-sized variable face_mask";
    return false;
  }
  att_double(r, "time", s.time);
  att_double(r, "Omega", s.omega);
  int vp = -2, vf = -1, pv = 0, av = 0;
  att_int(r, "velocity_prec", vp);
  att_int(r, "velocity_form", vf);
  att_int(r, "streamfn_valid", pv);
  att_int(r, "absvort_valid", av);
  if (vp > -2 && r.has_var("velocityVerts")) {
    if (!need("velocityVerts", s.vu, 3 * s.nv) || !need("velocityFaces", s.fu, 3 * s.nf)) return false;
    s.vel_prec = vp;
    s.vel_form = vf;
  }
  if (pv && r.has_var("streamFnVerts")) {
    if (!need("streamFnVerts", s.vpsi, s.nv) || !need("streamFnFaces", s.fpsi, s.nf)) return false;
    s.psi_valid = true;
  }
  if (av && r.has_var("absVortVerts") && att_double(r, "absvort_omega", s.absv_omega)) {
    if (!need("absVortVerts", s.vabsv, s.nv) || !need("absVortFaces", s.fabsv, s.nf)) return false;
    s.absv_valid = true;
  }
  return true;
}

namespace {

// Text or binary values of a VTK file: legacy binary is big-endian; XML
// binary is base64 of a UInt64 byte count followed by the little-endian data.
struct VtkOut {
  std::ofstream& f;
  bool binary, xml;

  static void base64(const std::vector<uint8_t>& b, std::string& out) {
    static const char* t = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    out.clear();
    out.reserve((b.size() + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < b.size(); i += 3) {
      const uint32_t v = (uint32_t)b[i] << 16 | (uint32_t)b[i + 1] << 8 | b[i + 2];
      out += t[v >> 18];
      out += t[(v >> 12) & 63];
      out += t[(v >> 6) & 63];
      out += t[v & 63];
    }
    if (i < b.size()) {
      const uint32_t v = (uint32_t)b[i] << 16 | (i + 1 < b.size() ? (uint32_t)b[i + 1] << 8 : 0);
      out += t[v >> 18];
      out += t[(v >> 12) & 63];
      out += i + 1 < b.size() ? t[(v >> 6) & 63] : '=';
      out += '=';
    }
  }
  template <typename T>
  void values(const T* p, int64_t n, int per_line, const char* fmt) {
    if (binary && xml) {
      std::vector<uint8_t> b(8 + sizeof(T) * (size_t)n);
      const uint64_t nbytes = sizeof(T) * (uint64_t)n;
      std::memcpy(b.data(), &nbytes, 8);  // the host is little-endian
      std::memcpy(b.data() + 8, p, (size_t)nbytes);
      std::string s;
      base64(b, s);
      f << s << '\n';
      return;
    }
    if (binary) {
      std::vector<uint8_t> b(sizeof(T) * (size_t)n);
      for (int64_t i = 0; i < n; ++i) {
        const uint8_t* c = (const uint8_t*)&p[i];
        for (size_t k = 0; k < sizeof(T); ++k) b[sizeof(T) * (size_t)i + k] = c[sizeof(T) - 1 - k];
      }
      f.write((const char*)b.data(), (std::streamsize)b.size());
      f << '\n';
      return;
    }
    char buf[40];
    for (int64_t i = 0; i < n; ++i) {
      std::snprintf(buf, sizeof buf, fmt, p[i]);
      f << buf << (((i + 1) % per_line == 0) ? '\n' : ' ');
    }
    if (n % per_line) f << '\n';
  }
  void doubles(const double* p, int64_t n, int per_line) { values(p, n, per_line, "%.17g"); }
};

struct Array {
  const char* name;
  int ncomp;
  const double* data;  // n tuples
};

// Array names as the VTK writers store them: legacy files write blanks,
// control characters, '%' and non-ASCII bytes as %XX (the names are
// blank-separated tokens there); XML attributes escape the markup characters.
std::string legacy_name(const std::string& n) {
  std::string out;
  char buf[4];
  for (unsigned char ch : n) {
    if (ch <= ' ' || ch > '~' || ch == '%') {
      std::snprintf(buf, sizeof buf, "%%%02X", ch);
      out += buf;
    } else {
      out += (char)ch;
    }
  }
  return out;
}

std::string xml_name(const std::string& n) {
  std::string out;
  for (char ch : n) {
    switch (ch) {
      case '&': out += "&amp;"; break;
      case '<': out += "&lt;"; break;
      case '>': out += "&gt;"; break;
      case '"': out += "&quot;"; break;
      case '\'': out += "&apos;"; break;
      default: out += ch;
    }
  }
  return out;
}

}  // namespace

namespace {

// Leaf faces of the tree (no kids, src/LpmVtkIO.cpp:23-30).
std::vector<int64_t> leaf_faces(const TreeMesh& m) {
  std::vector<int64_t> leaves;
  for (int64_t f = 0; f < m.nfaces(); ++f)
    if (m.face_kids[(size_t)(4 * f)] < 0) leaves.push_back(f);
  return leaves;
}

// Polydata of points vx[nv][3] and the listed leaf faces, with point and cell
// arrays (cell arrays already restricted to the leaves).
bool write_polydata(const std::string& path, int format, int64_t nv, const double* vx, const TreeMesh& m,
                    const std::vector<int64_t>& leaves, const std::vector<Array>& pdata,
                    const std::vector<Array>& cdata, const char* title, std::string& err) {
  if (format < kVtkLegacyAscii || format > kVtkXmlBinary) {
    err = "unknown VTK format";
    return false;
  }
  const bool xml = format >= kVtkXmlAscii, binary = format == kVtkLegacyBinary || format == kVtkXmlBinary;
  const int64_t nc = (int64_t)leaves.size();
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    err = "cannot open " + path + " for writing";
    return false;
  }
  VtkOut o{f, binary, xml};
  if (!xml) {
    f << "# vtk DataFile Version 3.0\n" << title << '\n' << (binary ? "BINARY" : "ASCII")
      << "\nDATASET POLYDATA\nPOINTS " << nv << " double\n";
    o.doubles(vx, 3 * nv, 3);
    const int nfv = m.nfaceverts;
    f << "POLYGONS " << nc << ' ' << (nfv + 1) * nc << '\n';
    std::vector<int32_t> conn((size_t)((nfv + 1) * nc));
    for (int64_t c = 0; c < nc; ++c) {
      conn[(size_t)((nfv + 1) * c)] = nfv;
      for (int k = 0; k < nfv; ++k)
        conn[(size_t)((nfv + 1) * c + 1 + k)] = m.face_verts[(size_t)(nfv * leaves[(size_t)c] + k)];
    }
    o.values(conn.data(), (int64_t)conn.size(), nfv + 1, "%d");
    auto fields = [&](const std::vector<Array>& as, int64_t n) {
      f << "FIELD FieldData " << as.size() << '\n';
      for (const Array& a : as) {
        f << legacy_name(a.name) << ' ' << a.ncomp << ' ' << n << " double\n";
        o.doubles(a.data, n * a.ncomp, a.ncomp == 1 ? 9 : a.ncomp);
      }
    };
    if (!pdata.empty()) {
      f << "POINT_DATA " << nv << '\n';
      fields(pdata, nv);
    }
    if (!cdata.empty()) {
      f << "CELL_DATA " << nc << '\n';
      fields(cdata, nc);
    }
  } else {
    const char* fmt = binary ? "binary" : "ascii";
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" "
         "header_type=\"UInt64\">\n<PolyData>\n<Piece NumberOfPoints=\""
      << nv << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"" << nc
      << "\">\n";
    auto arrays = [&](const char* tag, const std::vector<Array>& as, int64_t n) {
      f << '<' << tag << ">\n";
      for (const Array& a : as) {
        f << "<DataArray type=\"Float64\" Name=\"" << xml_name(a.name) << "\" NumberOfComponents=\"" << a.ncomp
          << "\" format=\"" << fmt << "\">\n";
        o.doubles(a.data, n * a.ncomp, a.ncomp == 1 ? 9 : a.ncomp);
        f << "</DataArray>\n";
      }
      f << "</" << tag << ">\n";
    };
    arrays("PointData", pdata, nv);
    arrays("CellData", cdata, nc);
    f << "<Points>\n<DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"" << fmt
      << "\">\n";
    o.doubles(vx, 3 * nv, 3);
    f << "</DataArray>\n</Points>\n<Polys>\n";
    const int nfv = m.nfaceverts;
    std::vector<long long> conn((size_t)(nfv * nc)), offs((size_t)nc);
    for (int64_t c = 0; c < nc; ++c) {
      for (int k = 0; k < nfv; ++k) conn[(size_t)(nfv * c + k)] = m.face_verts[(size_t)(nfv * leaves[(size_t)c] + k)];
      offs[(size_t)c] = nfv * (c + 1);
    }
    f << "<DataArray type=\"Int64\" Name=\"connectivity\" format=\"" << fmt << "\">\n";
    o.values(conn.data(), (int64_t)conn.size(), 9, "%lld");
    f << "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"" << fmt << "\">\n";
    o.values(offs.data(), (int64_t)offs.size(), 9, "%lld");
    f << "</DataArray>\n</Polys>\n</Piece>\n</PolyData>\n</VTKFile>\n";
  }
  if (!f) {
    err = "write failed: " + path;
    return false;
  }
  return true;
}

// Caller arrays after a writer's own: point arrays as given, face arrays
// (tree order) restricted to the leaf rows, kept alive in `leaf_rows`.
void add_extras(const std::vector<VtkExtra>& extra, const std::vector<int64_t>& leaves, std::vector<Array>& pdata,
                std::vector<Array>& cdata, std::vector<std::vector<double>>& leaf_rows) {
  const int64_t nc = (int64_t)leaves.size();
  leaf_rows.reserve(extra.size());
  for (const VtkExtra& x : extra) {
    if (!x.faces) {
      pdata.push_back({x.name.c_str(), x.ncomp, x.data});
      continue;
    }
    leaf_rows.emplace_back((size_t)(x.ncomp * nc));
    std::vector<double>& rows = leaf_rows.back();
    for (int64_t c = 0; c < nc; ++c)
      for (int k = 0; k < x.ncomp; ++k) rows[(size_t)(x.ncomp * c + k)] = x.data[x.ncomp * leaves[(size_t)c] + k];
    cdata.push_back({x.name.c_str(), x.ncomp, rows.data()});
  }
}

}  // namespace

bool write_bve_vtk(const std::string& path, const BveFields& s, const TreeMesh& m, int format, std::string& err,
                   const std::vector<VtkExtra>& extra) {
  if (m.nverts() != s.nv || m.nfaces() != s.nf) {
    err = "mesh sizes differ from the state's";
    return false;
  }
  const bool xml = format >= kVtkXmlAscii;
  const std::vector<int64_t> leaves = leaf_faces(m);
  const int64_t nc = (int64_t)leaves.size();
  std::vector<double> c_area(nc), c_zeta(nc), c_absv(nc), c_psi(nc), c_u(3 * nc);
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t f = leaves[(size_t)c];
    c_area[c] = s.farea[f];
    c_zeta[c] = s.fzeta[f];
    c_absv[c] = s.fabsv[f];
    c_psi[c] = s.fpsi[f];
    for (int k = 0; k < 3; ++k) c_u[3 * c + k] = s.fu[3 * f + k];
  }
  const char* n_zeta = xml ? "relvort" : "relVort";
  const char* n_absv = xml ? "absvort" : "absVort";
  const char* n_psi = xml ? "stream_fn" : "streamFn";
  std::vector<Array> pdata = {{n_zeta, 1, s.vzeta.data()},
                              {n_absv, 1, s.vabsv.data()},
                              {n_psi, 1, s.vpsi.data()},
                              {"velocity", 3, s.vu.data()}};
  std::vector<Array> cdata = {{"area", 1, c_area.data()},
                              {n_zeta, 1, c_zeta.data()},
                              {n_absv, 1, c_absv.data()},
                              {n_psi, 1, c_psi.data()},
                              {"velocity", 3, c_u.data()}};
  std::vector<std::vector<double>> extra_leaf;  // leaf rows of face arrays
  add_extras(extra, leaves, pdata, cdata, extra_leaf);
  return write_polydata(path, format, s.nv, s.vx.data(), m, leaves, pdata, cdata, "LPM BVE sphere", err);
}

bool vtk_extras(const lpm_vtk_array* arrays, int narrays, std::vector<VtkExtra>& out, std::string& err) {
  out.clear();
  if (narrays < 0 || (narrays > 0 && !arrays)) {
    err = "bad array list";
    return false;
  }
  for (int k = 0; k < narrays; ++k) {
    const lpm_vtk_array& a = arrays[k];
    if (!a.name || !a.name[0] || std::strlen(a.name) > 255 || a.ncomp < 1 || a.ncomp > 9 || !a.data) {
      err = "array " + std::to_string(k) + ": name must have 1 to 255 characters, 1 <= ncomp <= 9, data non-null";
      return false;
    }
    out.push_back({a.name, a.ncomp, a.on_faces != 0, a.data});
  }
  return true;
}

bool write_mesh_vtk(const std::string& path, const TreeMesh& m, int format, std::string& err,
                    const std::vector<VtkExtra>& extra) {
  const std::vector<int64_t> leaves = leaf_faces(m);
  std::vector<double> area(leaves.size());
  for (size_t c = 0; c < leaves.size(); ++c) area[c] = m.face_area[(size_t)leaves[c]];
  std::vector<Array> pdata, cdata = {{"area", 1, area.data()}};
  std::vector<std::vector<double>> extra_leaf;
  add_extras(extra, leaves, pdata, cdata, extra_leaf);
  return write_polydata(path, format, m.nverts(), m.vert_phys.data()->v, m, leaves, pdata, cdata, "LPM PolyMesh2d",
                        err);
}

}  // namespace lpm
