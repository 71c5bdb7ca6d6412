// Tree meshes on the unit sphere, built on the host: the icosahedral
// triangle seed and the cubed-sphere quadrilateral seed.
//
// Mirrors MeshSeed<IcosTriSphereSeed> / MeshSeed<CubedSphereSeed>
// (src/LpmMeshSeed.hpp:111-147, :181-224) and PolyMesh2d<Seed>::treeInit
// (src/LpmPolyMesh2d.cpp:30-44): every leaf face is split 4 ways, level by
// level, with edge midpoints projected to the sphere (a quadrilateral's
// centre becomes its children's shared corner).  Vertex, face and edge
// numbering, positions and areas are identical to the reference's, so
// particle i here is particle i there.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lpm_geometry.hpp"

namespace lpm {

struct IcosTriSphereSeed {
  static constexpr int nverts = 12;
  static constexpr int nfaces = 20;
  static constexpr int nedges = 30;
  static constexpr int nfaceverts = 3;
  static const char* filename() { return "icosTriSphereSeed.dat"; }
  static const char* idString() { return "IcosTriSphereSeed"; }
  // Counts at tree level `lev` (src/LpmMeshSeed.cpp:381-391).
  static int64_t nVerticesAtTreeLevel(int lev) { return 2 + 10 * (int64_t(1) << (2 * lev)); }
  static int64_t nFacesAtTreeLevel(int lev) { return 20 * (int64_t(1) << (2 * lev)); }
  static int64_t nEdgesAtTreeLevel(int64_t nv, int64_t nf) { return nv + nf - 2; }
};

// The two spherical seeds (lpm_seed_kind of lpm_bve.h).
enum SeedKind { kSeedIcosTri = 0, kSeedCubedSphere = 1 };

struct CubedSphereSeed {
  static constexpr int nverts = 8;
  static constexpr int nfaces = 6;
  static constexpr int nedges = 12;
  static constexpr int nfaceverts = 4;
  static const char* filename() { return "cubedSphereSeed.dat"; }
  static const char* idString() { return "CubedSphereSeed"; }
  // Counts at tree level `lev` (src/LpmMeshSeed.cpp:369-379).
  static int64_t nVerticesAtTreeLevel(int lev) { return 2 + 6 * (int64_t(1) << (2 * lev)); }
  static int64_t nFacesAtTreeLevel(int lev) { return 6 * (int64_t(1) << (2 * lev)); }
  static int64_t nEdgesAtTreeLevel(int64_t nv, int64_t nf) { return nf + nv - 2; }
};

// Seed data: 12 vertices then 20 face centres, 30 edges (origin, destination,
// left face, right face), 20 faces (vertices, edges).
struct SeedTables {
  Vec3 crds[32];
  int edges[30][4];
  int faceverts[20][3];
  int faceedges[20][3];
};

// The icosahedral seed shipped with the library (the values of the
// reference's mesh_seeds/icosTriSphereSeed.dat, needed bit for bit for
// identical meshes).
const SeedTables& builtin_icos_seed();

// Parse a seed file in the reference's format (src/LpmMeshSeed.cpp:20-232).
// Returns false and sets `err` on failure.
bool read_icos_seed(const std::string& path, SeedTables& out, std::string& err);

// Allocation sizes for a tree of depth `lev` (MeshSeed::setMaxAllocations,
// src/LpmMeshSeed.cpp:296-305): vertices, edges, faces (all tree levels).
void icos_max_allocations(int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces);
void cubed_max_allocations(int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces);
// by seed kind; false for an unknown kind
bool seed_max_allocations(int kind, int lev, int64_t& nverts, int64_t& nedges, int64_t& nfaces);

// Cubed-sphere seed data: 8 vertices then 6 face centres, 12 edges, 6 faces
// (4 vertices, 4 edges each).
struct QuadSeedTables {
  Vec3 crds[14];
  int edges[12][4];
  int faceverts[6][4];
  int faceedges[6][4];
};

// The cubed-sphere seed shipped with the library (the values of the
// reference's mesh_seeds/cubedSphereSeed.dat).
const QuadSeedTables& builtin_cubed_seed();

// Parse a cubed-sphere seed file in the reference's format.
bool read_cubed_seed(const std::string& path, QuadSeedTables& out, std::string& err);

// The built mesh.  Faces include every tree level (parents first, leaves
// last); divided faces have area 0 and mask 1.
struct TreeMesh {
  int level = 0;
  int kind = kSeedIcosTri;
  int nfaceverts = 3;
  std::vector<Vec3> vert_phys, vert_lag;
  std::vector<Vec3> face_phys, face_lag;
  std::vector<double> face_area;
  std::vector<uint8_t> face_mask;
  std::vector<int> face_verts;   // [nfaces][nfaceverts]
  std::vector<int> face_parent;  // -1 for seed faces
  int64_t n_edges = 0;
  // topology of Edges and Faces (src/LpmEdges.hpp, src/LpmFaces.hpp); -1 is
  // the reference's NULL_IND (src/LpmDefs.hpp:102)
  std::vector<int> edge_orig, edge_dest, edge_left, edge_right, edge_parent;
  std::vector<int> edge_kids;   // [nedges][2]
  std::vector<int> face_edges;  // [nfaces][nfaceverts]
  std::vector<int> face_kids;   // [nfaces][4]
  std::vector<int> face_level;  // tree level, 0 for seed faces
  int64_t nverts() const { return (int64_t)vert_phys.size(); }
  int64_t nfaces() const { return (int64_t)face_phys.size(); }
  int64_t nleaves() const;
  const char* seed_id() const { return kind == kSeedCubedSphere ? "CubedSphereSeed" : "IcosTriSphereSeed"; }
  const char* face_kind() const { return nfaceverts == 4 ? "quad" : "tri"; }
};

// PolyMesh2d::treeInit: seed, then `lev` rounds of 4-way refinement.
TreeMesh build_icos_tree(int lev, const SeedTables& seed);
TreeMesh build_cubed_tree(int lev, const QuadSeedTables& seed);
// the tree of a seed kind from its built-in seed
TreeMesh build_tree(int kind, int lev);

}  // namespace lpm
