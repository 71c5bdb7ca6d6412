// Host mesh build: icosahedral and cubed-sphere seeds and 4-way tree
// refinement.  Restates src/LpmMeshSeed.cpp:20-232,296-320,369-391,
// src/LpmPolyMesh2d.cpp:10-44, src/LpmFaces.cpp:9-63,115-381 and
// src/LpmEdges.cpp:12-85 (see lpm_mesh.hpp).  Compiled with -ffp-contract=off.
#include "lpm_mesh.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace lpm {

namespace {

constexpr int kNone = -1;

const SeedTables kIcosSeed = {
    {
        {{0.0, 0.0, 1.0}},
        {{0.7236067977499789, 0.5257311121191336, 0.4472135954999579}},
        {{-0.276393202250021, 0.8506508083520399, 0.4472135954999579}},
        {{-0.8944271909999159, 0.0, 0.4472135954999579}},
        {{-0.276393202250021, -0.8506508083520399, 0.4472135954999579}},
        {{0.7236067977499789, -0.5257311121191336, 0.4472135954999579}},
        {{0.8944271909999159, 0.0, -0.4472135954999579}},
        {{0.276393202250021, 0.8506508083520399, -0.4472135954999579}},
        {{-0.7236067977499789, 0.5257311121191336, -0.4472135954999579}},
        {{-0.7236067977499789, -0.5257311121191336, -0.4472135954999579}},
        {{0.276393202250021, -0.8506508083520399, -0.4472135954999579}},
        {{0.0, 0.0, -1.0}},
        {{0.18759247408507992, 0.5773502691896258, 0.7946544722917662}},
        {{-0.491123473188423, 0.35682208977308993, 0.7946544722917661}},
        {{-0.491123473188423, -0.35682208977308993, 0.7946544722917661}},
        {{0.18759247408507992, -0.5773502691896258, 0.7946544722917662}},
        {{0.6070619982066863, 0.0, 0.7946544722917662}},
        {{0.7946544722917661, 0.5773502691896257, -0.1875924740850799}},
        {{0.30353099910334314, 0.9341723589627157, 0.18759247408507992}},
        {{-0.30353099910334314, 0.9341723589627157, -0.18759247408507992}},
        {{-0.7946544722917661, 0.5773502691896257, 0.1875924740850799}},
        {{-0.982246946376846, 0.0, -0.1875924740850799}},
        {{-0.7946544722917661, -0.5773502691896257, 0.1875924740850799}},
        {{-0.30353099910334314, -0.9341723589627157, -0.18759247408507992}},
        {{0.30353099910334314, -0.9341723589627157, 0.18759247408507992}},
        {{0.7946544722917661, -0.5773502691896257, -0.1875924740850799}},
        {{0.982246946376846, 0.0, 0.1875924740850799}},
        {{0.491123473188423, 0.35682208977308993, -0.7946544722917661}},
        {{-0.18759247408507992, 0.5773502691896258, -0.7946544722917662}},
        {{-0.6070619982066863, 0.0, -0.7946544722917662}},
        {{-0.18759247408507992, -0.5773502691896258, -0.7946544722917662}},
        {{0.491123473188423, -0.35682208977308993, -0.7946544722917661}},
    },
    {{0, 1, 0, 4}, {1, 2, 0, 6}, {2, 0, 0, 1}, {2, 3, 1, 8}, {0, 3, 2, 1}, {3, 4, 2, 10}, {4, 0, 2, 3}, {4, 5, 3, 12}, {5, 0, 3, 4}, {5, 1, 4, 14}, {1, 6, 5, 14}, {6, 7, 5, 15}, {7, 1, 5, 6}, {7, 2, 6, 7}, {7, 8, 7, 16}, {8, 2, 7, 8}, {8, 3, 8, 9}, {8, 9, 9, 17}, {3, 9, 10, 9}, {9, 4, 10, 11}, {9, 10, 11, 18}, {10, 4, 11, 12}, {10, 5, 12, 13}, {10, 6, 13, 19}, {6, 5, 13, 14}, {11, 6, 19, 15}, {11, 7, 15, 16}, {8, 11, 17, 16}, {11, 9, 17, 18}, {10, 11, 19, 18}},
    {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5}, {0, 5, 1}, {1, 6, 7}, {7, 2, 1}, {2, 7, 8}, {8, 3, 2}, {3, 8, 9}, {9, 4, 3}, {4, 9, 10}, {10, 5, 4}, {5, 10, 6}, {6, 1, 5}, {11, 7, 6}, {11, 8, 7}, {11, 9, 8}, {11, 10, 9}, {11, 6, 10}},
    {{0, 1, 2}, {2, 3, 4}, {4, 5, 6}, {6, 7, 8}, {8, 9, 0}, {10, 11, 12}, {13, 1, 12}, {13, 14, 15}, {16, 3, 15}, {16, 17, 18}, {19, 5, 18}, {19, 20, 21}, {22, 7, 21}, {22, 23, 24}, {10, 9, 24}, {26, 11, 25}, {27, 14, 26}, {28, 17, 27}, {29, 20, 28}, {25, 23, 29}},
};

const double kC = 0.57735026918962584;  // 1/sqrt(3), as the seed file prints it
const QuadSeedTables kCubedSeed = {
    {
        {{kC, -kC, kC}},
        {{kC, -kC, -kC}},
        {{kC, kC, -kC}},
        {{kC, kC, kC}},
        {{-kC, kC, -kC}},
        {{-kC, kC, kC}},
        {{-kC, -kC, -kC}},
        {{-kC, -kC, kC}},
        {{1.0, 0.0, 0.0}},
        {{0.0, 1.0, 0.0}},
        {{-1.0, 0.0, 0.0}},
        {{0.0, -1.0, 0.0}},
        {{0.0, 0.0, 1.0}},
        {{0.0, 0.0, -1.0}},
    },
    {{0, 1, 0, 3}, {1, 2, 0, 5}, {2, 3, 0, 1}, {3, 0, 0, 4}, {2, 4, 1, 5}, {4, 5, 1, 2},
     {5, 3, 1, 4}, {4, 6, 2, 5}, {6, 7, 2, 3}, {7, 5, 2, 4}, {6, 1, 3, 5}, {0, 7, 3, 4}},
    {{0, 1, 2, 3}, {3, 2, 4, 5}, {5, 4, 6, 7}, {7, 6, 1, 0}, {7, 0, 3, 5}, {1, 6, 4, 2}},
    {{0, 1, 2, 3}, {2, 4, 5, 6}, {5, 7, 8, 9}, {8, 10, 0, 11}, {11, 3, 6, 9}, {10, 7, 4, 1}},
};

// Growing edge and face tables of the tree.  Edges store origin,
// destination, left and right face, and the two children of a divided
// edge (child 0 starts at the origin).
struct Builder {
  TreeMesh& m;
  const int nfv;  // vertices (and edges) per face
  std::vector<int> e_orig, e_dest, e_left, e_right, e_kid0, e_kid1;
  std::vector<int> f_edges;  // [nfaces][nfv]
  std::vector<int> f_kid0;   // first child of a divided face, else -1

  explicit Builder(TreeMesh& mesh) : m(mesh), nfv(mesh.nfaceverts) {}

  int add_edge(int o, int d, int l, int r) {
    e_orig.push_back(o);
    e_dest.push_back(d);
    e_left.push_back(l);
    e_right.push_back(r);
    e_kid0.push_back(kNone);
    e_kid1.push_back(kNone);
    return (int)e_orig.size() - 1;
  }

  int add_vertex(const Vec3& phys, const Vec3& lag) {
    m.vert_phys.push_back(phys);
    m.vert_lag.push_back(lag);
    return (int)m.vert_phys.size() - 1;
  }

  // Split edge e at its projected midpoint; children inherit both faces.
  void split_edge(int e) {
    const int o = e_orig[e], d = e_dest[e], l = e_left[e], r = e_right[e];
    const int mid = add_vertex(sphere::midpoint(m.vert_phys[o], m.vert_phys[d]),
                               sphere::midpoint(m.vert_lag[o], m.vert_lag[d]));
    const int c0 = add_edge(o, mid, l, r);
    const int c1 = add_edge(mid, d, l, r);
    e_kid0[e] = c0;
    e_kid1[e] = c1;
  }

  void add_face(const int* v, const int* e, int parent, const Vec3& ctr, const Vec3& lagctr, double area) {
    for (int k = 0; k < nfv; ++k) {
      m.face_verts.push_back(v[k]);
      f_edges.push_back(e[k]);
    }
    f_kid0.push_back(kNone);
    m.face_parent.push_back(parent);
    m.face_phys.push_back(ctr);
    m.face_lag.push_back(lagctr);
    m.face_area.push_back(area);
    m.face_mask.push_back(0);
  }

  // Split face f into corner children 0..2 (child k keeps corner k) and the
  // centre child 3.
  void split_face(int f) {
    const int first = (int)m.face_phys.size();
    int cv[4][3], ce[4][3];
    for (auto& row : cv) row[0] = row[1] = row[2] = kNone;
    for (auto& row : ce) row[0] = row[1] = row[2] = kNone;
    for (int k = 0; k < 3; ++k) cv[k][k] = m.face_verts[3 * f + k];
    for (int k = 0; k < 3; ++k) {
      const int pe = f_edges[3 * f + k];
      if (!(e_kid0[pe] > 0)) split_edge(pe);
      const int h0 = e_kid0[pe], h1 = e_kid1[pe];
      const int next = (k + 1) % 3;
      if (e_left[pe] == f) {  // edge runs counter-clockwise around f
        ce[k][k] = h0;
        e_left[h0] = first + k;
        ce[next][k] = h1;
        e_left[h1] = first + next;
      } else {
        ce[k][k] = h1;
        e_right[h1] = first + k;
        ce[next][k] = h0;
        e_right[h0] = first + next;
      }
      const int mid = e_dest[h0];
      // midpoint of parent edge k is shared by corner children k and k+1
      // and by the centre child (opposite corner k+2)
      cv[k][next] = mid;
      cv[next][k] = mid;
      cv[3][(k + 2) % 3] = mid;
    }
    const int e0 = (int)e_orig.size();
    ce[3][0] = e0;
    ce[3][1] = e0 + 1;
    ce[3][2] = e0 + 2;
    ce[0][1] = e0 + 1;
    ce[1][2] = e0 + 2;
    ce[2][0] = e0;
    add_edge(cv[2][1], cv[2][0], first + 3, first + 2);
    add_edge(cv[0][2], cv[0][1], first + 3, first + 0);
    add_edge(cv[1][0], cv[1][2], first + 3, first + 1);
    for (int c = 0; c < 4; ++c) {
      Vec3 phys[3], lag[3];
      for (int k = 0; k < 3; ++k) {
        phys[k] = m.vert_phys[cv[c][k]];
        lag[k] = m.vert_lag[cv[c][k]];
      }
      const Vec3 ctr = sphere::barycenter(phys, 3);
      const Vec3 lagctr = sphere::barycenter(lag, 3);
      add_face(cv[c], ce[c], f, ctr, lagctr, sphere::polygon_area(ctr, phys, 3));
    }
    f_kid0[f] = first;
    m.face_area[f] = 0.0;
    m.face_mask[f] = 1;
  }

  // FaceDivider<SphereGeometry,QuadFace>::divide (src/LpmFaces.cpp:251-381):
  // child k keeps corner k; the midpoint of parent edge k is the corner
  // shared by children k and k+1; the parent's centre becomes a vertex,
  // the corner all four children share (their corner k+2); four interior
  // edges join it to the midpoints.
  void split_quad(int f) {
    const int first = (int)m.face_phys.size();
    int cv[4][4], ce[4][4];
    for (auto& row : cv) row[0] = row[1] = row[2] = row[3] = kNone;
    for (auto& row : ce) row[0] = row[1] = row[2] = row[3] = kNone;
    for (int k = 0; k < 4; ++k) cv[k][k] = m.face_verts[4 * f + k];
    for (int k = 0; k < 4; ++k) {
      const int pe = f_edges[4 * f + k];
      if (!(e_kid0[pe] > 0)) split_edge(pe);
      const int h0 = e_kid0[pe], h1 = e_kid1[pe];
      const int next = (k + 1) % 4;
      if (e_left[pe] == f) {
        ce[k][k] = h0;
        e_left[h0] = first + k;
        ce[next][k] = h1;
        e_left[h1] = first + next;
      } else {
        ce[k][k] = h1;
        e_right[h1] = first + k;
        ce[next][k] = h0;
        e_right[h0] = first + next;
      }
      const int mid = e_dest[h0];
      cv[k][next] = mid;
      cv[next][k] = mid;
    }
    // the parent's centre particle (face f's coordinates) as the shared corner
    const int centre = add_vertex(m.face_phys[f], m.face_lag[f]);
    for (int k = 0; k < 4; ++k) cv[k][(k + 2) % 4] = centre;
    const int e0 = (int)e_orig.size();
    add_edge(cv[0][1], cv[0][2], first + 0, first + 1);
    ce[0][1] = e0;
    ce[1][3] = e0;
    add_edge(cv[2][0], cv[2][3], first + 3, first + 2);
    ce[2][3] = e0 + 1;
    ce[3][1] = e0 + 1;
    add_edge(cv[2][1], cv[2][0], first + 1, first + 2);
    ce[1][2] = e0 + 2;
    ce[2][0] = e0 + 2;
    add_edge(cv[3][1], cv[3][0], first + 0, first + 3);
    ce[0][2] = e0 + 3;
    ce[3][0] = e0 + 3;
    for (int c = 0; c < 4; ++c) {
      Vec3 phys[4], lag[4];
      for (int k = 0; k < 4; ++k) {
        phys[k] = m.vert_phys[cv[c][k]];
        lag[k] = m.vert_lag[cv[c][k]];
      }
      const Vec3 ctr = sphere::barycenter(phys, 4);
      const Vec3 lagctr = sphere::barycenter(lag, 4);
      add_face(cv[c], ce[c], f, ctr, lagctr, sphere::polygon_area(ctr, phys, 4));
    }
    f_kid0[f] = first;
    m.face_area[f] = 0.0;
    m.face_mask[f] = 1;
  }
};

// Line-oriented reader for the reference's seed format: a header line, the
// vertex and face-centre coordinates, then the "edgeO", "faceverts" and
// "faceedges" sections (src/LpmMeshSeed.cpp:43-230).
bool read_seed_file(const std::string& path, int nverts, int nfaces, int nedges, int nfv, Vec3* crds, int* edges,
                    int* faceverts, int* faceedges, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open seed file " + path;
    return false;
  }
  const int ncrd = nverts + nfaces;
  int edge_hdr = -1, fv_hdr = -1, fe_hdr = -1, n_crd = 0, n_e = 0, n_fv = 0, n_fe = 0, line_no = 0;
  std::string line;
  while (std::getline(in, line)) {
    ++line_no;
    if (line.find("edgeO") != std::string::npos) edge_hdr = line_no;
    if (line.find("faceverts") != std::string::npos) fv_hdr = line_no;
    if (line.find("faceedges") != std::string::npos) fe_hdr = line_no;
    std::istringstream ss(line);
    bool ok = true;
    if (line_no > 1 && line_no < ncrd + 2) {
      Vec3& p = crds[n_crd++];
      ok = bool(ss >> p[0] >> p[1] >> p[2]);
    } else if (edge_hdr > 0 && line_no > edge_hdr && line_no <= edge_hdr + nedges) {
      int* e = edges + 4 * n_e++;
      ok = bool(ss >> e[0] >> e[1] >> e[2] >> e[3]);
    } else if (fv_hdr > 0 && line_no > fv_hdr && line_no <= fv_hdr + nfaces) {
      int* v = faceverts + nfv * n_fv++;
      for (int k = 0; k < nfv && ok; ++k) ok = bool(ss >> v[k]);
    } else if (fe_hdr > 0 && line_no > fe_hdr && line_no <= fe_hdr + nfaces) {
      int* e = faceedges + nfv * n_fe++;
      for (int k = 0; k < nfv && ok; ++k) ok = bool(ss >> e[k]);
    }
    if (!ok) {
      err = "malformed line " + std::to_string(line_no) + " in " + path;
      return false;
    }
  }
  if (n_crd != ncrd || n_e != nedges || n_fv != nfaces || n_fe != nfaces) {
    err = "incomplete seed file " + path;
    return false;
  }
  return true;
}

// Tree construction shared by both seeds: seed vertices, edges and faces
// (a seed face keeps the seed file's centre, which its area uses too), then
// `lev` rounds of dividing every leaf in index order, then the topology
// arrays.
template <typename SeedT, typename Tables>
TreeMesh build_tree_from(int lev, const Tables& seed, int kind) {
  TreeMesh m;
  m.level = lev;
  m.kind = kind;
  m.nfaceverts = SeedT::nfaceverts;
  constexpr int nfv = SeedT::nfaceverts;
  int64_t nv = SeedT::nVerticesAtTreeLevel(lev), ne = 0, nf = 0;
  for (int i = 0; i <= lev; ++i) {
    nf += SeedT::nFacesAtTreeLevel(i);
    ne += SeedT::nEdgesAtTreeLevel(SeedT::nVerticesAtTreeLevel(i), SeedT::nFacesAtTreeLevel(i));
  }
  m.vert_phys.reserve(nv);
  m.vert_lag.reserve(nv);
  m.face_phys.reserve(nf);
  m.face_lag.reserve(nf);
  m.face_area.reserve(nf);
  m.face_mask.reserve(nf);
  m.face_verts.reserve(nfv * nf);
  m.face_parent.reserve(nf);
  Builder b(m);
  b.e_orig.reserve(ne);
  b.e_dest.reserve(ne);
  b.e_left.reserve(ne);
  b.e_right.reserve(ne);
  b.e_kid0.reserve(ne);
  b.e_kid1.reserve(ne);
  b.f_edges.reserve(nfv * nf);
  b.f_kid0.reserve(nf);
  for (int i = 0; i < SeedT::nverts; ++i) b.add_vertex(seed.crds[i], seed.crds[i]);
  for (int i = 0; i < SeedT::nedges; ++i)
    b.add_edge(seed.edges[i][0], seed.edges[i][1], seed.edges[i][2], seed.edges[i][3]);
  for (int i = 0; i < SeedT::nfaces; ++i) {
    const Vec3& ctr = seed.crds[SeedT::nverts + i];
    Vec3 corners[nfv];
    for (int k = 0; k < nfv; ++k) corners[k] = seed.crds[seed.faceverts[i][k]];
    b.add_face(seed.faceverts[i], seed.faceedges[i], kNone, ctr, ctr, sphere::polygon_area(ctr, corners, nfv));
  }
  for (int l = 0; l < lev; ++l) {
    const int stop = (int)m.face_phys.size();
    for (int f = 0; f < stop; ++f) {
      if (b.f_kid0[f] > 0) continue;
      if (nfv == 4)
        b.split_quad(f);
      else
        b.split_face(f);
    }
  }
  m.n_edges = (int64_t)b.e_orig.size();
  // topology: edges with parents derived from their kids; faces with kids
  // (consecutive, first at f_kid0) and tree levels
  const int64_t ne_all = m.n_edges, nf_all = m.nfaces();
  m.edge_orig = b.e_orig;
  m.edge_dest = b.e_dest;
  m.edge_left = b.e_left;
  m.edge_right = b.e_right;
  m.edge_parent.assign(ne_all, kNone);
  m.edge_kids.assign(2 * ne_all, kNone);
  for (int64_t e = 0; e < ne_all; ++e) {
    m.edge_kids[2 * e] = b.e_kid0[e];
    m.edge_kids[2 * e + 1] = b.e_kid1[e];
    if (b.e_kid0[e] != kNone) {
      m.edge_parent[b.e_kid0[e]] = (int)e;
      m.edge_parent[b.e_kid1[e]] = (int)e;
    }
  }
  m.face_edges = b.f_edges;
  m.face_kids.assign(4 * nf_all, kNone);
  m.face_level.assign(nf_all, 0);
  for (int64_t f = 0; f < nf_all; ++f) {
    if (m.face_parent[f] != kNone) m.face_level[f] = m.face_level[m.face_parent[f]] + 1;
    if (b.f_kid0[f] != kNone)
      for (int k = 0; k < 4; ++k) m.face_kids[4 * f + k] = b.f_kid0[f] + k;
  }
  return m;
}

}  // namespace

const SeedTables& builtin_icos_seed() { return kIcosSeed; }

bool read_icos_seed(const std::string& path, SeedTables& out, std::string& err) {
  return read_seed_file(path, IcosTriSphereSeed::nverts, IcosTriSphereSeed::nfaces, IcosTriSphereSeed::nedges,
                        IcosTriSphereSeed::nfaceverts, out.crds, &out.edges[0][0], &out.faceverts[0][0],
                        &out.faceedges[0][0], err);
}

const QuadSeedTables& builtin_cubed_seed() { return kCubedSeed; }

bool read_cubed_seed(const std::string& path, QuadSeedTables& out, std::string& err) {
  return read_seed_file(path, CubedSphereSeed::nverts, CubedSphereSeed::nfaces, CubedSphereSeed::nedges,
                        CubedSphereSeed::nfaceverts, out.crds, &out.edges[0][0], &out.faceverts[0][0],
                        &out.faceedges[0][0], err);
}

void icos_max_allocations(int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces) {
  nverts = IcosTriSphereSeed::nVerticesAtTreeLevel(lev);
  nedges = 0;
  nfaces = 0;
  for (int i = 0; i <= lev; ++i) {
    nfaces += IcosTriSphereSeed::nFacesAtTreeLevel(i);
    nedges += IcosTriSphereSeed::nEdgesAtTreeLevel(IcosTriSphereSeed::nVerticesAtTreeLevel(i),
                                                   IcosTriSphereSeed::nFacesAtTreeLevel(i));
  }
}

void cubed_max_allocations(int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces) {
  nverts = CubedSphereSeed::nVerticesAtTreeLevel(lev);
  nedges = 0;
  nfaces = 0;
  for (int i = 0; i <= lev; ++i) {
    nfaces += CubedSphereSeed::nFacesAtTreeLevel(i);
    nedges += CubedSphereSeed::nEdgesAtTreeLevel(CubedSphereSeed::nVerticesAtTreeLevel(i),
                                                 CubedSphereSeed::nFacesAtTreeLevel(i));
  }
}

bool seed_max_allocations(int kind, int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces) {
  if (kind == kSeedIcosTri)
    icos_max_allocations(lev, nverts, nedges, nfaces);
  else if (kind == kSeedCubedSphere)
    cubed_max_allocations(lev, nverts, nedges, nfaces);
  else
    return false;
  return true;
}

int64_t TreeMesh::nleaves() const {
  int64_t n = 0;
  for (uint8_t k : face_mask) n += (k == 0);
  return n;
}

TreeMesh build_icos_tree(int lev, const SeedTables& seed) {
  return build_tree_from<IcosTriSphereSeed>(lev, seed, kSeedIcosTri);
}

TreeMesh build_cubed_tree(int lev, const QuadSeedTables& seed) {
  return build_tree_from<CubedSphereSeed>(lev, seed, kSeedCubedSphere);
}

TreeMesh build_tree(int kind, int lev) {
  return kind == kSeedCubedSphere ? build_cubed_tree(lev, builtin_cubed_seed()) : build_icos_tree(lev, builtin_icos_seed());
}

}  // namespace lpm
