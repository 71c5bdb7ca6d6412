// Mesh and field files.
//
// * netCDF: the PolyMesh2d variables of NcWriter::writePolymesh
//   (src/LpmNetCDF_Impl.hpp:46-218) and the BVESphere fields under their view
//   labels (src/LpmBVESphere.cpp:13-20), as NcWriter::writeScalarField /
//   writeVectorField (src/LpmNetCDF_Impl.hpp:222-296) name them.  The same
//   file carries what resuming a run needs (time, rotation rate, which
//   velocity and absolute vorticity the state holds).
// * VTK: legacy polydata as BVESphere::outputVtk (src/LpmBVESphere.cpp:90-113)
//   writes it, and XML polydata (.vtp) as Polymesh2dVtkInterface::write with
//   BVESphere::addFieldsToVtk (src/LpmBVESphere.cpp:171-187,
//   src/LpmPolyMesh2dVtkInterface_Impl.hpp:136-141).  Cells are the leaf faces.
#pragma once

#include <string>
#include <vector>

#include "lpm_bve.h"
#include "lpm_mesh.hpp"
#include "lpm_ncio.hpp"

namespace lpm {

// Adds the dimensions, variables and attributes of writePolymesh for `m` to
// `w`.  phys_v [nverts][3] / phys_f [nfaces][3] replace the mesh's physical
// coordinates when non-null (the particles' current positions).  Pointers
// into `m` and the arguments must stay valid until w.write().
// phys_v, phys_f, area (may be null: the tree's own) replace the tree's
// physical coordinates and face areas.
void add_mesh_vars(nc::Writer& w, const TreeMesh& m, const double* phys_v, const double* phys_f,
                   const double* area = nullptr);

// A BVE state on the host, faces in tree order (divided faces included).
struct BveFields {
  int64_t nv = 0, nf = 0;
  std::vector<double> vx, vzeta, vabsv, vpsi, vu;  // [nv][3] for vx, vu
  std::vector<double> fx, fzeta, fabsv, fpsi, fu;  // [nf][3] for fx, fu
  std::vector<double> farea;
  std::vector<uint8_t> fmask;
  double time = 0.0, omega = 0.0;
  // velocity: -2 none, -1 given by the caller, else the precision and pair
  // form of the sum that produced it
  int vel_prec = -2, vel_form = -1;
  bool psi_valid = false;
  bool absv_valid = false;
  double absv_omega = 0.0;
};

// Writes `s` to a netCDF file: with `mesh` (the tree the particles were built
// on) every writePolymesh variable, else the particle variables only
// (phys_crds_verts, phys_crds_faces, face_area, face_mask); then the BVE
// fields and the resume attributes.
bool write_bve_nc(const std::string& path, const BveFields& s, const TreeMesh* mesh, std::string& err);

// Reads a file written by write_bve_nc (or a writePolymesh file plus the
// BVE field variables).  Missing relVort* is an error; missing velocity or
// absolute vorticity leaves vel_prec = -2 / absv_valid = false.
bool read_bve_nc(const std::string& path, BveFields& s, std::string& err);

enum VtkFormat { kVtkLegacyAscii = 0, kVtkLegacyBinary = 1, kVtkXmlAscii = 2, kVtkXmlBinary = 3 };

// A caller's array added after the state's own (as the reference's drivers add
// error fields with addVectorPointData / addVectorCellData): [nverts][ncomp]
// at vertices, or [nfaces][ncomp] in tree order at faces (leaf rows written).
struct VtkExtra {
  std::string name;
  int ncomp = 1;
  bool faces = false;
  const double* data = nullptr;
};

// Writes `s` on `mesh`'s leaf faces.  Legacy formats use outputVtk's array
// names (relVort, absVort, streamFn, velocity; cell "area" first); XML
// formats those of addFieldsToVtk (relvort, absvort, stream_fn, velocity).
bool write_bve_vtk(const std::string& path, const BveFields& s, const TreeMesh& mesh, int format,
                   std::string& err, const std::vector<VtkExtra>& extra = {});

// The caller arrays of the C ABI as VtkExtra: false (err set) unless every
// name has 1 to 255 bytes, 1 <= ncomp <= 9 and data is non-null.
bool vtk_extras(const lpm_vtk_array* arrays, int narrays, std::vector<VtkExtra>& out, std::string& err);

// PolyMesh2d::outputVtk (src/LpmPolyMesh2d.cpp:79-85): the mesh's leaf faces
// with their areas as cell data ("area"), then `extra` arrays (the fields
// SpherePoisson::outputVtk adds, src/LpmSpherePoisson.hpp:510-531).
bool write_mesh_vtk(const std::string& path, const TreeMesh& mesh, int format, std::string& err,
                    const std::vector<VtkExtra>& extra = {});

}  // namespace lpm
