// lpm_bve.hpp -- C++ host interface over the C ABI of lpm_bve.h, in the
// shape of the reference's BVE classes:
//
//   MeshSeed<IcosTriSphereSeed> seed;  (or CubedSphereSeed) src/LpmMeshSeed.hpp:181-224
//   seed.setMaxAllocations(nv, ne, nf, depth);             src/LpmMeshSeed.cpp:296-305
//   BVESphere<IcosTriSphereSeed> sphere(nv, ne, nf);        src/LpmBVESphere.hpp:40
//   sphere.treeInit(depth, seed);                           src/LpmPolyMesh2d.hpp:171
//   sphere.set_omega(omega);                                src/LpmBVESphere.cpp:115-124
//   sphere.init_vorticity(ic);                              src/LpmBVESphere.cpp:126-169
//   BVERK4 solver(dt, Omega); solver.init(nv, nf);          src/LpmBVERK4.hpp:29-34
//   solver.advance_timestep(sphere.physVerts.crds, sphere.relVortVerts, sphere.velocityVerts,
//                           sphere.physFaces.crds, sphere.relVortFaces, sphere.velocityFaces,
//                           sphere.faces.area, sphere.faces.mask);
//
// Fields live in device memory (one lpm_bve_ctx per sphere); the *View members
// name them, updateHost() fills the host mirrors.  VorticityInitialCondition
// is the reference's plugin interface (src/LpmVorticityGallery.hpp:13-23):
// any subclass is evaluated per particle on the host.
#ifndef LPM_BVE_HPP
#define LPM_BVE_HPP

#include <algorithm>
#include <cmath>
#include <iostream>
#include <cstdint>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpm_bve.h"

namespace Lpm {

typedef double Real;
typedef int64_t Index;
typedef short Short;
static constexpr Real PI = 3.1415926535897932384626433832795027975;

inline void check(int rc, const char* what) {
  if (rc != LPM_OK) throw std::runtime_error(std::string(what) + ": " + lpm_last_error());
}

// The spherical seeds (src/LpmMeshSeed.hpp:111-147).
struct IcosTriSphereSeed {
  static constexpr int nverts = 12, nfaces = 20, nedges = 30, nfaceverts = 3;
  static constexpr int kind = LPM_SEED_ICOS_TRI;
  static std::string idString() { return "IcosTriSphereSeed"; }
  static std::string faceStr() { return "tri"; }
};

struct CubedSphereSeed {
  static constexpr int nverts = 8, nfaces = 6, nedges = 12, nfaceverts = 4;
  static constexpr int kind = LPM_SEED_CUBED_SPHERE;
  static std::string idString() { return "CubedSphereSeed"; }
  static std::string faceStr() { return "quad"; }
};

template <typename SeedType>
struct MeshSeed {
  std::string path;  // empty: the library's built-in seed
  MeshSeed() {}
  explicit MeshSeed(const std::string& seed_file) : path(seed_file) {}
  static std::string idString() { return SeedType::idString(); }
  void setMaxAllocations(Index& nboundary, Index& nedges, Index& nfaces, const int lev) const {
    check(lpm_seed_alloc(SeedType::kind, lev, &nboundary, &nedges, &nfaces), "setMaxAllocations");
  }
  // PolyMesh2d<SeedType>::treeInit arrays (fverts [nfaces][nfaceverts]).
  void treeInit(int lev, double* vx, double* vlag, double* fx, double* flag, double* farea, uint8_t* fmask,
                int64_t* counts) const {
    check(lpm_seed_tree_init(SeedType::kind, lev, path.empty() ? nullptr : path.c_str(), vx, vlag, fx, flag, farea,
                             fmask, nullptr, nullptr, counts),
          "treeInit");
  }
};

struct VorticityInitialCondition {
  typedef std::shared_ptr<VorticityInitialCondition> ptr;
  virtual ~VorticityInitialCondition() {}
  virtual Real eval(const Real& x, const Real& y, const Real& z) const = 0;
  virtual std::string name() const { return std::string(); }
  // Evaluate at n points; the gallery fields below route this to the library.
  virtual void eval_many(const Real* xyz, Index n, Real* zeta) const {
    for (Index i = 0; i < n; ++i) zeta[i] = eval(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  }
};

// Composite gallery field evaluated by lpm_ic_eval: solid-body rotation,
// Rossby-Haurwitz wave, Gaussian vortices.
struct GalleryVorticity : public VorticityInitialCondition {
  Real sbr = 0, rh_w = 0, rh_K = 0, sigma = 1;
  int rh_n = 0;
  std::vector<Real> centers, amps;
  Real eval(const Real& x, const Real& y, const Real& z) const override {
    const Real p[3] = {x, y, z};
    Real out;
    eval_many(p, 1, &out);
    return out;
  }
  void eval_many(const Real* xyz, Index n, Real* zeta) const override {
    check(lpm_ic_eval(xyz, n, zeta, sbr, rh_n, rh_w, rh_K, (int)amps.size(), centers.data(), amps.data(), sigma),
          "lpm_ic_eval");
  }
  void add_vortices(uint64_t seed, int nvort, Real amp_scale, Real sig) {
    centers.assign(3 * nvort, 0.0);
    amps.assign(nvort, 0.0);
    check(lpm_ic_gauss_params(seed, nvort, amp_scale, centers.data(), amps.data()), "lpm_ic_gauss_params");
    sigma = sig;
  }
};

struct SolidBodyRotation : public GalleryVorticity {
  static constexpr Real OMEGA = 2 * PI;
  SolidBodyRotation() { sbr = 1.0; }
  std::string name() const override { return "rotation"; }
};

struct RossbyHaurwitzWave : public GalleryVorticity {
  RossbyHaurwitzWave(int n, Real w, Real K) {
    rh_n = n;
    rh_w = w;
    rh_K = K;
  }
  std::string name() const override { return "rossby_haurwitz"; }
  Real phase_speed(Real omega) const { return (rh_n * (rh_n + 3) * rh_w - 2 * omega) / ((rh_n + 1) * (rh_n + 2)); }
};

// ErrNorms (src/LpmErrorNorms.hpp:81-130, src/LpmErrorNorms.cpp:7-70):
// relative l1, l2 and linf errors with weights (areas), computed on the host
// by lpm_err_norms_vec / lpm_err_norms_scalar in index order.
struct ErrNorms {
  Real l1 = 0, l2 = 0, linf = 0;
  ErrNorms(const Real l_1, const Real l_2, const Real l_i) : l1(l_1), l2(l_2), linf(l_i) {}
  // vector fields [n][3]: err (resized) receives appx - exact
  ErrNorms(std::vector<Real>& err, const std::vector<Real>& appx, const std::vector<Real>& exact,
           const std::vector<Real>& wt) {
    const Index n = (Index)wt.size();
    if ((Index)appx.size() < 3 * n || (Index)exact.size() < 3 * n)
      throw std::invalid_argument("ErrNorms: vector fields need 3 values per weight");
    err.resize(3 * n);
    Real out[3];
    check(lpm_err_norms_vec(err.data(), appx.data(), exact.data(), wt.data(), n, out), "ErrNorms");
    l1 = out[0];
    l2 = out[1];
    linf = out[2];
  }
  // a scalar error field err = appx - exact
  ErrNorms(const std::vector<Real>& err, const std::vector<Real>& exact, const std::vector<Real>& wt) {
    const Index n = (Index)wt.size();
    if ((Index)err.size() < n || (Index)exact.size() < n)
      throw std::invalid_argument("ErrNorms: scalar fields need one value per weight");
    Real out[3];
    check(lpm_err_norms_scalar(err.data(), exact.data(), wt.data(), n, out), "ErrNorms");
    l1 = out[0];
    l2 = out[1];
    linf = out[2];
  }
  std::string infoString(const std::string& label = "", const int tab_level = 0) const {
    std::ostringstream ss;
    for (int i = 0; i < tab_level; ++i) ss << "\t";
    ss << label << " ErrNorms: l1 = " << l1 << " l2 = " << l2 << " linf = " << linf << "\n";
    return ss.str();
  }
};

// SphereVelocityTangentTestFunctor (src/LpmSphereTestKernels.hpp:11-25):
// udotx[i] = x_i . u_i for positions and velocities [n][3].
inline void velocity_tangent(std::vector<Real>& udotx, const std::vector<Real>& x, const std::vector<Real>& u) {
  const Index n = (Index)std::min(x.size(), u.size()) / 3;
  udotx.resize(n);
  check(lpm_velocity_tangent(x.data(), u.data(), n, udotx.data()), "velocity_tangent");
}

template <typename SeedType>
class BVESphere;

// Names one field of a sphere's device-resident state.
template <typename SeedType>
struct FieldView {
  BVESphere<SeedType>* sphere;
  int id;
};

template <typename SeedType>
class BVESphere {
 public:
  enum { VX, VZETA, VU, FX, FZETA, FU, FAREA, FMASK };
  struct CoordsView {
    FieldView<SeedType> crds;
  };
  struct FacesView {
    FieldView<SeedType> area, mask;
  };

  CoordsView physVerts, physFaces;
  FacesView faces;
  FieldView<SeedType> relVortVerts, relVortFaces, velocityVerts, velocityFaces;
  Real Omega = 2 * PI;
  Real t = 0;

  // passive tracers (src/LpmBVESphere.cpp:36-58): host arrays [nmaxverts] and
  // [nmaxfaces], written with the state by outputVtk
  std::vector<std::string> tracer_names;
  std::vector<std::vector<Real>> tracer_verts, tracer_faces;

  // host mirrors (filled by treeInit / init_vorticity / updateHost)
  std::vector<Real> vert_x, vert_lag, face_x, face_lag, face_area;
  std::vector<uint8_t> face_mask;
  std::vector<Real> rel_vort_verts, rel_vort_faces, abs_vort_verts, abs_vort_faces;
  std::vector<Real> vel_verts, vel_faces, psi_verts, psi_faces;

  BVESphere(Index nmaxverts, Index nmaxedges, Index nmaxfaces, int nq = 0, int device = 0)
      : nmaxv_(nmaxverts), nmaxf_(nmaxfaces) {
    (void)nmaxedges;
    physVerts.crds = {this, VX};
    physFaces.crds = {this, FX};
    faces.area = {this, FAREA};
    faces.mask = {this, FMASK};
    relVortVerts = {this, VZETA};
    relVortFaces = {this, FZETA};
    velocityVerts = {this, VU};
    velocityFaces = {this, FU};
    check(lpm_bve_create(device, &ctx_), "lpm_bve_create");
    for (int k = 0; k < nq; ++k) create_tracer("tracer" + std::to_string(k));
  }

  // BVESphere::create_tracer: a new zero tracer; returns its index.
  Short create_tracer(const std::string& name) {
    tracer_names.push_back(name);
    tracer_verts.emplace_back(nmaxv_, 0.0);
    tracer_faces.emplace_back(nmaxf_, 0.0);
    return (Short)(tracer_names.size() - 1);
  }
  ~BVESphere() { lpm_bve_destroy(ctx_); }
  BVESphere(const BVESphere&) = delete;
  BVESphere& operator=(const BVESphere&) = delete;

  void treeInit(int depth, const MeshSeed<SeedType>& seed) {
    Index nv, ne, nf;
    seed.setMaxAllocations(nv, ne, nf, depth);
    if (nv > nmaxv_ || nf > nmaxf_) throw std::runtime_error("treeInit: not enough memory");
    vert_x.assign(3 * nv, 0);
    vert_lag.assign(3 * nv, 0);
    face_x.assign(3 * nf, 0);
    face_lag.assign(3 * nf, 0);
    face_area.assign(nf, 0);
    face_mask.assign(nf, 0);
    int64_t counts[4];
    seed.treeInit(depth, vert_x.data(), vert_lag.data(), face_x.data(), face_lag.data(), face_area.data(),
                  face_mask.data(), counts);
    nverts_ = counts[0];
    nfaces_ = counts[2];
    nleaves_ = counts[3];
    depth_ = depth;
  }

  Index nvertsHost() const { return nverts_; }
  Index nfacesHost() const { return nfaces_; }
  Index nleavesHost() const { return nleaves_; }

  void set_omega(const Real& omg) {
    if (omg_set_) return;
    Omega = omg;
    omg_set_ = true;
  }

  // Host IC evaluation, absolute vorticity, upload, then the device solves
  // for stream function and velocity (src/LpmBVESphere.cpp:126-169).
  void init_vorticity(const VorticityInitialCondition::ptr relvort, int prec = LPM_PREC_FP64,
                      int scheme = LPM_SCHEME_REF) {
    rel_vort_verts.assign(nverts_, 0);
    rel_vort_faces.assign(nfaces_, 0);
    relvort->eval_many(vert_x.data(), nverts_, rel_vort_verts.data());
    relvort->eval_many(face_x.data(), nfaces_, rel_vort_faces.data());
    abs_vort_verts.resize(nverts_);
    abs_vort_faces.resize(nfaces_);
    for (Index i = 0; i < nverts_; ++i) abs_vort_verts[i] = rel_vort_verts[i] + 2 * Omega * vert_x[3 * i + 2];
    for (Index i = 0; i < nfaces_; ++i) abs_vort_faces[i] = rel_vort_faces[i] + 2 * Omega * face_x[3 * i + 2];
    check(lpm_bve_set_state(ctx_, nverts_, vert_x.data(), rel_vort_verts.data(), nfaces_, face_x.data(),
                            rel_vort_faces.data(), face_area.data(), face_mask.data()),
          "lpm_bve_set_state");
    check(lpm_bve_solve(ctx_, prec, scheme, 1), "lpm_bve_solve");
  }

  // Stream function at the current state (the post-step solve of
  // tests/LpmBVETest.cpp:155-159); the velocity, the next step's entry
  // velocity, is left as it is.
  void solve_streamfn(int prec = LPM_PREC_FP64, int scheme = LPM_SCHEME_REF) {
    check(lpm_bve_streamfn(ctx_, prec, scheme), "lpm_bve_streamfn");
  }

  void updateHost() {
    vel_verts.resize(3 * nverts_);
    vel_faces.resize(3 * nfaces_);
    psi_verts.resize(nverts_);
    psi_faces.resize(nfaces_);
    check(lpm_bve_get_state(ctx_, vert_x.data(), rel_vort_verts.data(), vel_verts.data(), psi_verts.data(),
                            face_x.data(), rel_vort_faces.data(), vel_faces.data(), psi_faces.data()),
          "lpm_bve_get_state");
  }

  // One more field for outputVtk: [nverts][ncomp] at vertices, or
  // [nfaces][ncomp] at faces in tree order.
  struct VtkArray {
    std::string name;
    int ncomp = 1;
    bool on_faces = false;
    std::vector<Real> data;
  };

  // BVESphere::outputVtk (src/LpmBVESphere.cpp:90-113): legacy VTK polydata of
  // the leaf faces; a name ending in .vtp writes the XML layout of
  // Polymesh2dVtkInterface::write with addFieldsToVtk.  The tracers follow the
  // state's fields, and `extra` fields follow them.
  void outputVtk(const std::string& fname, const std::vector<VtkArray>& extra = {}) {
    const bool vtp = fname.size() > 4 && fname.compare(fname.size() - 4, 4, ".vtp") == 0;
    std::vector<lpm_vtk_array> a;
    for (size_t k = 0; k < tracer_names.size(); ++k) {
      a.push_back({tracer_names[k].c_str(), 1, 0, tracer_verts[k].data()});
      a.push_back({tracer_names[k].c_str(), 1, 1, tracer_faces[k].data()});
    }
    for (const VtkArray& x : extra) a.push_back({x.name.c_str(), x.ncomp, x.on_faces ? 1 : 0, x.data.data()});
    check(lpm_bve_write_vtk_arrays(ctx_, fname.c_str(), depth_, Omega, vtp ? LPM_VTK_XML_BINARY : LPM_VTK_LEGACY_ASCII,
                                   a.data(), (int)a.size()),
          "outputVtk");
  }

  // The state, time and rotation rate in one netCDF file (the mesh variables of
  // NcWriter::writePolymesh and the fields under the view labels); restore()
  // continues bit for bit.
  void checkpoint(const std::string& fname) {
    check(lpm_bve_checkpoint(ctx_, fname.c_str(), depth_, t, Omega), "checkpoint");
  }
  void restore(const std::string& fname) {
    check(lpm_bve_restore(ctx_, fname.c_str(), &t, &Omega), "restore");
    omg_set_ = true;
    int64_t sz[4];
    check(lpm_bve_sizes(ctx_, sz), "lpm_bve_sizes");
    nverts_ = sz[0];
    nfaces_ = sz[1];
    nleaves_ = sz[2];
    depth_ = -1;
    const int64_t per_level = SeedType::nfaces / 2;  // nverts = 2 + (nfaces/2) 4^level on both seeds
    for (int lev = 0; lev <= 12; ++lev)
      if (2 + per_level * (int64_t(1) << (2 * lev)) == nverts_) depth_ = lev;
    vert_x.assign(3 * nverts_, 0);
    face_x.assign(3 * nfaces_, 0);
    rel_vort_verts.assign(nverts_, 0);
    rel_vort_faces.assign(nfaces_, 0);
  }

  Real avg_mesh_size_radians() const { return std::sqrt(4 * PI / nfaces_); }
  Real avg_mesh_size_degrees() const { return avg_mesh_size_radians() * 180.0 / PI; }
  Real circulation() {
    Real c;
    check(lpm_bve_circulation(ctx_, &c), "lpm_bve_circulation");
    return c;
  }
  lpm_bve_ctx* context() { return ctx_; }

 private:
  lpm_bve_ctx* ctx_ = nullptr;
  Index nmaxv_, nmaxf_;
  Index nverts_ = 0, nfaces_ = 0, nleaves_ = 0;
  int depth_ = -1;
  bool omg_set_ = false;
};

class BVERK4 {
 public:
  Real dt, Omega;
  Index nverts = 0, nfaces = 0;
  int scheme = LPM_SCHEME_REF, prec = LPM_PREC_FP64;
  bool use_graph = true;

  BVERK4(const Real& timestep, const Real& omg) : dt(timestep), Omega(omg) {}
  void init(const Index& nv, const Index& nf) {
    nverts = nv;
    nfaces = nf;
  }

  template <typename SeedType>
  void advance_timestep(FieldView<SeedType> vx, FieldView<SeedType> vzeta, FieldView<SeedType> vvel,
                        FieldView<SeedType> fx, FieldView<SeedType> fzeta, FieldView<SeedType> fvel,
                        FieldView<SeedType> fa, FieldView<SeedType> fm, int nsteps = 1) {
    typedef BVESphere<SeedType> S;
    BVESphere<SeedType>* s = vx.sphere;
    const FieldView<SeedType> v[8] = {vx, vzeta, vvel, fx, fzeta, fvel, fa, fm};
    const int ids[8] = {S::VX, S::VZETA, S::VU, S::FX, S::FZETA, S::FU, S::FAREA, S::FMASK};
    for (int k = 0; k < 8; ++k)
      if (v[k].sphere != s || v[k].id != ids[k])
        throw std::invalid_argument("advance_timestep: views must be one sphere's fields, in order");
    if (s->nvertsHost() != nverts || s->nfacesHost() != nfaces)
      throw std::invalid_argument("advance_timestep: init(nverts, nfaces) does not match the sphere");
    check(lpm_bve_advance(s->context(), dt, Omega, scheme, prec, nsteps, use_graph ? 1 : 0), "lpm_bve_advance");
    s->t += nsteps * dt;
  }
};

// SpherePoisson<Seed> (src/LpmSpherePoisson.hpp:359-520): -lap(psi) = f on
// the sphere with f = Y_5^4 (exact psi = f/30, exact velocity RH54Velocity),
// solved by the same direct sums as the BVE path (stream function and
// velocity at vertices and faces, FP64, literal 1 - x.y, on the GPU), with the
// reference's L-infinity errors over vertices and active faces.
template <typename SeedType = IcosTriSphereSeed>
class SpherePoisson {
 public:
  std::vector<Real> fverts, psiverts, psiexactverts, ffaces, psifaces, psiexactfaces;
  std::vector<Real> everts, efaces;
  std::vector<Real> uverts, ufaces, uvertsexact, ufacesexact, uerrverts, uerrfaces;
  Real linf_verts = 0, linf_faces = 0, linf_uverts = 0, linf_ufaces = 0;
  std::vector<Real> vert_x, face_x, face_area;
  std::vector<uint8_t> face_mask;

  SpherePoisson(Index nmaxverts, Index nmaxedges, Index nmaxfaces, int device = 0)
      : nmaxv_(nmaxverts), nmaxf_(nmaxfaces), device_(device) {
    (void)nmaxedges;
  }

  void treeInit(int depth, const MeshSeed<SeedType>& seed) {
    Index nv, ne, nf;
    seed.setMaxAllocations(nv, ne, nf, depth);
    if (nv > nmaxv_ || nf > nmaxf_) throw std::runtime_error("treeInit: not enough memory");
    vert_x.assign(3 * nv, 0);
    face_x.assign(3 * nf, 0);
    face_area.assign(nf, 0);
    face_mask.assign(nf, 0);
    int64_t counts[4];
    seed.treeInit(depth, vert_x.data(), nullptr, face_x.data(), nullptr, face_area.data(), face_mask.data(), counts);
    nverts_ = counts[0];
    nfaces_ = counts[2];
    nleaves_ = counts[3];
    depth_ = depth;
    seed_kind_ = SeedType::kind;
    seed_path_ = seed.path;
  }

  Index nvertsHost() const { return nverts_; }
  Index nfacesHost() const { return nfaces_; }

  // SpherePoisson::outputVtk (src/LpmSpherePoisson.hpp:510-531): legacy VTK
  // polydata of the leaf faces (XML for a name ending in .vtp) with the
  // right-hand side, solution, exact solution and their errors at vertices
  // and faces, cell "area" first.
  void outputVtk(const std::string& fname) const {
    const bool vtp = fname.size() > 4 && fname.compare(fname.size() - 4, 4, ".vtp") == 0;
    const lpm_vtk_array a[] = {
        {"f", 1, 0, fverts.data()},          {"psi", 1, 0, psiverts.data()},
        {"psi_exact", 1, 0, psiexactverts.data()}, {"psi_error", 1, 0, everts.data()},
        {"velocity", 3, 0, uverts.data()},   {"velocity_exact", 3, 0, uvertsexact.data()},
        {"velocity_error", 3, 0, uerrverts.data()},
        {"f", 1, 1, ffaces.data()},          {"psi", 1, 1, psifaces.data()},
        {"psi_exact", 1, 1, psiexactfaces.data()}, {"psi_error", 1, 1, efaces.data()},
        {"velocity", 3, 1, ufaces.data()},   {"velocity_exact", 3, 1, ufacesexact.data()},
        {"velocity_error", 3, 1, uerrfaces.data()}};
    check(lpm_seed_mesh_write_vtk_arrays(seed_kind_, depth_, seed_path_.empty() ? nullptr : seed_path_.c_str(),
                                         fname.c_str(), vtp ? LPM_VTK_XML_BINARY : LPM_VTK_LEGACY_ASCII, a,
                                         (int)(sizeof(a) / sizeof(a[0]))),
          "outputVtk");
  }

  // average mesh size (src/LpmSpherePoisson.hpp:410-415)
  Real meshSize() const { return std::sqrt(4 * PI / nleaves_); }

  // right-hand side and exact solution at every particle (Init, :60-84)
  void init() {
    fverts.assign(nverts_, 0);
    ffaces.assign(nfaces_, 0);
    uvertsexact.assign(3 * nverts_, 0);
    ufacesexact.assign(3 * nfaces_, 0);
    check(lpm_poisson_y54(vert_x.data(), nverts_, fverts.data(), uvertsexact.data()), "lpm_poisson_y54");
    check(lpm_poisson_y54(face_x.data(), nfaces_, ffaces.data(), ufacesexact.data()), "lpm_poisson_y54");
    psiexactverts.resize(nverts_);
    psiexactfaces.resize(nfaces_);
    for (Index i = 0; i < nverts_; ++i) psiexactverts[i] = fverts[i] / 30.0;
    for (Index i = 0; i < nfaces_; ++i) psiexactfaces[i] = ffaces[i] / 30.0;
  }

  // Stream function and velocity at vertices and faces, then the errors
  // (:428-475); prints the reference's lines.
  void solve(const int& nthreads = 0) {
    (void)nthreads;
    lpm_bve_ctx* ctx = nullptr;
    check(lpm_bve_create(device_, &ctx), "lpm_bve_create");
    psiverts.assign(nverts_, 0);
    psifaces.assign(nfaces_, 0);
    uverts.assign(3 * nverts_, 0);
    ufaces.assign(3 * nfaces_, 0);
    int rc = lpm_bve_set_state(ctx, nverts_, vert_x.data(), fverts.data(), nfaces_, face_x.data(), ffaces.data(),
                               face_area.data(), face_mask.data());
    if (rc == LPM_OK) rc = lpm_bve_solve(ctx, LPM_PREC_FP64, LPM_SCHEME_REF, 1);
    if (rc == LPM_OK)
      rc = lpm_bve_get_state(ctx, nullptr, nullptr, uverts.data(), psiverts.data(), nullptr, nullptr, ufaces.data(),
                             psifaces.data());
    lpm_bve_destroy(ctx);
    check(rc, "SpherePoisson::solve");
    everts.assign(nverts_, 0);
    efaces.assign(nfaces_, 0);
    uerrverts.assign(3 * nverts_, 0);
    uerrfaces.assign(3 * nfaces_, 0);
    linf_verts = linf_faces = linf_uverts = linf_ufaces = 0;
    for (Index i = 0; i < nverts_; ++i) {
      everts[i] = std::abs(psiverts[i] - psiexactverts[i]);
      linf_verts = std::max(linf_verts, everts[i]);
      Real e2 = 0;
      for (int k = 0; k < 3; ++k) {
        uerrverts[3 * i + k] = std::abs(uverts[3 * i + k] - uvertsexact[3 * i + k]);
        e2 += uerrverts[3 * i + k] * uerrverts[3 * i + k];
      }
      linf_uverts = std::max(linf_uverts, e2);
    }
    for (Index i = 0; i < nfaces_; ++i) {
      if (face_mask[i]) continue;  // divided faces: error 0
      efaces[i] = std::abs(psifaces[i] - psiexactfaces[i]);
      linf_faces = std::max(linf_faces, efaces[i]);
      Real e2 = 0;
      for (int k = 0; k < 3; ++k) {
        uerrfaces[3 * i + k] = std::abs(ufaces[3 * i + k] - ufacesexact[3 * i + k]);
        e2 += uerrfaces[3 * i + k] * uerrfaces[3 * i + k];
      }
      linf_ufaces = std::max(linf_ufaces, e2);
    }
    std::cout << meshSize() * (180.0 / PI) << ":\n"
              << "\tpotential: linf_verts = " << linf_verts << ", linf_faces = " << linf_faces << '\n'
              << "\tvelocity:  linf_verts = " << std::sqrt(linf_uverts) << ", linf_faces = " << std::sqrt(linf_ufaces)
              << '\n';
  }

  void updateHost() const {}
  void updateDevice() const {}

 private:
  Index nmaxv_, nmaxf_;
  int device_;
  Index nverts_ = 0, nfaces_ = 0, nleaves_ = 0;
  int depth_ = 0, seed_kind_ = LPM_SEED_ICOS_TRI;
  std::string seed_path_;
};

}  // namespace Lpm

#endif  // LPM_BVE_HPP
