// C ABI for the host-side pieces: mesh build, initial conditions, errors.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <exception>
#include <memory>
#include <string>

#include "lpm_bve.h"
#include "lpm_ic.hpp"
#include "lpm_mesh.hpp"
#include "lpm_meshio.hpp"
#include "lpm_status.hpp"

namespace lpm {
namespace {
thread_local std::string g_last_error;
}
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int ok() {
  g_last_error.clear();
  return LPM_OK;
}
}  // namespace lpm

using namespace lpm;

extern "C" {

const char* lpm_last_error(void) { return g_last_error.c_str(); }

int lpm_version(void) { return 10000; }

}  // extern "C"

namespace {

// The tree of seed `kind` at `level` from the built-in seed or a seed file.
int make_tree(int kind, int level, const char* seed_path, const char* who, std::unique_ptr<TreeMesh>& out) {
  if (kind != kSeedIcosTri && kind != kSeedCubedSphere) return fail(LPM_ERR_ARG, std::string(who) + ": unknown seed kind");
  if (level < 0 || level > 12) return fail(LPM_ERR_ARG, std::string(who) + ": level must be in [0, 12]");
  try {
    std::string err;
    if (kind == kSeedCubedSphere) {
      QuadSeedTables file_seed;
      const QuadSeedTables* seed = &builtin_cubed_seed();
      if (seed_path) {
        if (!read_cubed_seed(seed_path, file_seed, err)) return fail(LPM_ERR_IO, std::string(who) + ": " + err);
        seed = &file_seed;
      }
      out.reset(new TreeMesh(build_cubed_tree(level, *seed)));
    } else {
      SeedTables file_seed;
      const SeedTables* seed = &builtin_icos_seed();
      if (seed_path) {
        if (!read_icos_seed(seed_path, file_seed, err)) return fail(LPM_ERR_IO, std::string(who) + ": " + err);
        seed = &file_seed;
      }
      out.reset(new TreeMesh(build_icos_tree(level, *seed)));
    }
  } catch (const std::exception& e) {
    return fail(LPM_ERR_NOMEM, std::string(who) + ": " + e.what());
  }
  return LPM_OK;
}

}  // namespace

extern "C" {

int lpm_seed_alloc(int kind, int level, int64_t* nverts, int64_t* nedges, int64_t* nfaces) {
  if (level < 0 || level > 12) return fail(LPM_ERR_ARG, "lpm_seed_alloc: level must be in [0, 12]");
  int64_t v, e, f;
  if (!seed_max_allocations(kind, level, v, e, f)) return fail(LPM_ERR_ARG, "lpm_seed_alloc: unknown seed kind");
  if (nverts) *nverts = v;
  if (nedges) *nedges = e;
  if (nfaces) *nfaces = f;
  return ok();
}

int lpm_icos_alloc(int level, int64_t* nverts, int64_t* nedges, int64_t* nfaces) {
  if (level < 0 || level > 12) return fail(LPM_ERR_ARG, "lpm_icos_alloc: level must be in [0, 12]");
  return lpm_seed_alloc(kSeedIcosTri, level, nverts, nedges, nfaces);
}

int lpm_seed_tree_init(int kind, int level, const char* seed_path, double* vx, double* vlag, double* fx, double* flag,
                       double* farea, uint8_t* fmask, int32_t* fverts, int32_t* fparent, int64_t* counts) {
  std::unique_ptr<TreeMesh> tm;
  const int rc = make_tree(kind, level, seed_path, "lpm_seed_tree_init", tm);
  if (rc) return rc;
  const TreeMesh& m = *tm;
  const int64_t nv = m.nverts(), nf = m.nfaces();
  if (vx) std::memcpy(vx, m.vert_phys.data(), sizeof(double) * 3 * nv);
  if (vlag) std::memcpy(vlag, m.vert_lag.data(), sizeof(double) * 3 * nv);
  if (fx) std::memcpy(fx, m.face_phys.data(), sizeof(double) * 3 * nf);
  if (flag) std::memcpy(flag, m.face_lag.data(), sizeof(double) * 3 * nf);
  if (farea) std::memcpy(farea, m.face_area.data(), sizeof(double) * nf);
  if (fmask) std::memcpy(fmask, m.face_mask.data(), nf);
  if (fverts)
    for (int64_t i = 0; i < m.nfaceverts * nf; ++i) fverts[i] = m.face_verts[i];
  if (fparent)
    for (int64_t i = 0; i < nf; ++i) fparent[i] = m.face_parent[i];
  if (counts) {
    counts[0] = nv;
    counts[1] = m.n_edges;
    counts[2] = nf;
    counts[3] = m.nleaves();
  }
  return ok();
}

int lpm_icos_tree_init(int level, const char* seed_path, double* vx, double* vlag, double* fx, double* flag,
                       double* farea, uint8_t* fmask, int32_t* fverts, int32_t* fparent, int64_t* counts) {
  return lpm_seed_tree_init(kSeedIcosTri, level, seed_path, vx, vlag, fx, flag, farea, fmask, fverts, fparent, counts);
}

int lpm_seed_mesh_write_nc(int kind, int level, const char* seed_path, const char* path) {
  if (!path) return fail(LPM_ERR_ARG, "lpm_mesh_write_nc: null path");
  std::unique_ptr<TreeMesh> tm;
  const int rc = make_tree(kind, level, seed_path, "lpm_mesh_write_nc", tm);
  if (rc) return rc;
  nc::Writer w;
  add_mesh_vars(w, *tm, nullptr, nullptr);
  std::string err;
  if (!w.write(path, err)) return fail(LPM_ERR_IO, "lpm_mesh_write_nc: " + err);
  return ok();
}

int lpm_mesh_write_nc(int level, const char* seed_path, const char* path) {
  return lpm_seed_mesh_write_nc(kSeedIcosTri, level, seed_path, path);
}

int lpm_seed_mesh_write_vtk(int kind, int level, const char* seed_path, const char* path, int format) {
  if (!path) return fail(LPM_ERR_ARG, "lpm_mesh_write_vtk: null path");
  if (format < LPM_VTK_LEGACY_ASCII || format > LPM_VTK_XML_BINARY)
    return fail(LPM_ERR_ARG, "lpm_mesh_write_vtk: unknown format");
  std::unique_ptr<TreeMesh> tm;
  const int rc = make_tree(kind, level, seed_path, "lpm_mesh_write_vtk", tm);
  if (rc) return rc;
  std::string err;
  if (!write_mesh_vtk(path, *tm, format, err)) return fail(LPM_ERR_IO, "lpm_mesh_write_vtk: " + err);
  return ok();
}

int lpm_seed_mesh_write_vtk_arrays(int kind, int level, const char* seed_path, const char* path, int format,
                                   const lpm_vtk_array* arrays, int narrays) {
  if (!path) return fail(LPM_ERR_ARG, "lpm_mesh_write_vtk: null path");
  if (format < LPM_VTK_LEGACY_ASCII || format > LPM_VTK_XML_BINARY)
    return fail(LPM_ERR_ARG, "lpm_mesh_write_vtk: unknown format");
  std::vector<VtkExtra> extra;
  std::string err;
  if (!vtk_extras(arrays, narrays, extra, err)) return fail(LPM_ERR_ARG, "lpm_mesh_write_vtk: " + err);
  std::unique_ptr<TreeMesh> tm;
  const int rc = make_tree(kind, level, seed_path, "lpm_mesh_write_vtk", tm);
  if (rc) return rc;
  if (!write_mesh_vtk(path, *tm, format, err, extra)) return fail(LPM_ERR_IO, "lpm_mesh_write_vtk: " + err);
  return ok();
}

int lpm_mesh_write_vtk(int level, const char* seed_path, const char* path, int format) {
  return lpm_seed_mesh_write_vtk(kSeedIcosTri, level, seed_path, path, format);
}

int lpm_mesh_read_nc(const char* path, double* vx, double* vlag, double* fx, double* flag, double* farea,
                     uint8_t* fmask, int32_t* fverts, int32_t* fparent, int64_t* counts) {
  if (!path) return fail(LPM_ERR_ARG, "lpm_mesh_read_nc: null path");
  nc::Reader r;
  std::string err;
  if (!r.open(path, err)) return fail(LPM_ERR_IO, "lpm_mesh_read_nc: " + err);
  const int64_t nv = r.dim("nverts"), ne = r.dim("nedges"), nf = r.dim("nfaces");
  if (nv < 0 || nf < 1 || r.dim("ndim") != 3)
    return fail(LPM_ERR_IO, "lpm_mesh_read_nc: missing or bad nverts / nfaces / ndim dimensions");
  std::vector<uint8_t> mask;
  if (!r.read("face_mask", mask) || (int64_t)mask.size() != nf)
    return fail(LPM_ERR_IO, "lpm_mesh_read_nc: missing or mis-sized variable face_mask");
  auto get_d = [&](const char* name, double* out, int64_t n) -> bool {
    if (!out) return true;
    std::vector<double> v;
    if (!r.read(name, v) || (int64_t)v.size() != n) {
      err = name;
      return false;
    }
    std::memcpy(out, v.data(), sizeof(double) * n);
    return true;
  };
  auto get_i = [&](const char* name, int32_t* out, int64_t n) -> bool {
    if (!out) return true;
    std::vector<int32_t> v;
    if (!r.read(name, v) || (int64_t)v.size() != n) {
      err = name;
      return false;
    }
    std::memcpy(out, v.data(), sizeof(int32_t) * n);
    return true;
  };
  if (!get_d("phys_crds_verts", vx, 3 * nv) || !get_d("lag_crds_verts", vlag, 3 * nv) ||
      !get_d("phys_crds_faces", fx, 3 * nf) || !get_d("lag_crds_faces", flag, 3 * nf) ||
      !get_d("face_area", farea, nf) || !get_i("face_verts", fverts, std::max<int64_t>(r.dim("nfaceverts"), 0) * nf) ||
      !get_i("face_parents", fparent, nf))
    return fail(LPM_ERR_IO, "lpm_mesh_read_nc: missing or mis-sized variable " + err);
  if (fmask) std::memcpy(fmask, mask.data(), nf);
  if (counts) {
    int64_t leaves = 0;
    for (uint8_t b : mask) leaves += b == 0;
    counts[0] = nv;
    counts[1] = ne;
    counts[2] = nf;
    counts[3] = leaves;
  }
  return ok();
}

int lpm_ic_gauss_params(uint64_t seed, int nvort, double amp_scale, double* centers, double* amps) {
  if (nvort < 0) return fail(LPM_ERR_ARG, "lpm_ic_gauss_params: nvort < 0");
  CompositeVorticity ic;
  ic.draw_vortices(seed, nvort, amp_scale, 1.0);
  for (int k = 0; k < nvort; ++k) {
    for (int j = 0; j < 3; ++j) centers[3 * k + j] = ic.centers[k][j];
    amps[k] = ic.amps[k];
  }
  return ok();
}

int lpm_rng_draws(uint64_t seed, int64_t nuniform, double* uniform, int64_t nnormal, double* normal) {
  if (nuniform < 0 || nnormal < 0 || (nuniform && !uniform) || (nnormal && !normal))
    return fail(LPM_ERR_ARG, "lpm_rng_draws: bad arguments");
  Xoshiro256ss rng(seed);
  for (int64_t i = 0; i < nuniform; ++i) uniform[i] = rng.uniform();
  for (int64_t i = 0; i < nnormal; ++i) normal[i] = rng.normal();
  return ok();
}

int lpm_ic_eval(const double* xyz, int64_t n, double* zeta, double sbr, int rh_n, double rh_w, double rh_K,
                int nvort, const double* centers, const double* amps, double sigma) {
  if (n < 0 || nvort < 0 || (nvort > 0 && (!centers || !amps)))
    return fail(LPM_ERR_ARG, "lpm_ic_eval: bad arguments");
  CompositeVorticity ic;
  ic.sbr = sbr;
  ic.rh_n = rh_n;
  ic.rh_w = rh_w;
  ic.rh_K = rh_K;
  ic.sigma = sigma;
  for (int k = 0; k < nvort; ++k) {
    ic.centers.push_back(Vec3{{centers[3 * k], centers[3 * k + 1], centers[3 * k + 2]}});
    ic.amps.push_back(amps[k]);
  }
  // per-particle evaluation, as BVESphere::init_vorticity (src/LpmBVESphere.cpp:126-149)
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) zeta[i] = ic.eval(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  return ok();
}

int lpm_poisson_y54(const double* xyz, int64_t n, double* f, double* u) {
  if (n < 0 || (n && !xyz)) return fail(LPM_ERR_ARG, "lpm_poisson_y54: bad arguments");
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const Vec3 p{{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}};
    if (f) f[i] = sph_harm54(p);
    if (u) {
      const Vec3 v = rh54_velocity(p);
      for (int k = 0; k < 3; ++k) u[3 * i + k] = v[k];
    }
  }
  return ok();
}

int lpm_area_mean(const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, double* mean) {
  double num = 0.0, den = 0.0;
  for (int64_t i = 0; i < nf; ++i) {
    if (fmask && fmask[i]) continue;
    num += fzeta[i] * farea[i];
    den += farea[i];
  }
  if (den == 0.0) return fail(LPM_ERR_ARG, "lpm_area_mean: zero total area");
  *mean = num / den;
  return ok();
}

// ErrNorms::compute for vector fields (src/LpmErrorNorms.cpp:15-45): the
// error appx - exact (written to err when non-null), then in index order
// l1 = sum |e| w / sum |x| w, l2 = sqrt(sum |e|^2 w / sum |x|^2 w) and
// linf = max |e| / max |x|, with |v| = sqrt(v.v) (SphereGeometry::mag).
int lpm_err_norms_vec(double* err, const double* appx, const double* exact, const double* wt, int64_t n,
                      double* out) {
  if (n < 1 || !appx || !exact || !wt || !out) return fail(LPM_ERR_ARG, "lpm_err_norms_vec: bad arguments");
  double l1n = 0, l1d = 0, l2n = 0, l2d = 0, lin = -DBL_MAX, lid = -DBL_MAX;
  for (int64_t i = 0; i < n; ++i) {
    const double e[3] = {appx[3 * i] - exact[3 * i], appx[3 * i + 1] - exact[3 * i + 1],
                         appx[3 * i + 2] - exact[3 * i + 2]};
    if (err)
      for (int k = 0; k < 3; ++k) err[3 * i + k] = e[k];
    const double* x = exact + 3 * i;
    const double em = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const double xm = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    l1n += em * wt[i];
    l1d += xm * wt[i];
    l2n += (em * em) * wt[i];
    l2d += (xm * xm) * wt[i];
    if (em > lin) lin = em;
    if (xm > lid) lid = xm;
  }
  out[0] = l1n / l1d;
  out[1] = std::sqrt(l2n / l2d);
  out[2] = lin / lid;
  return ok();
}

// ErrNorms::compute for a scalar error field err = appx - exact
// (src/LpmErrorNorms.cpp:56-70): l1 and l2 as above with |e| and |x|; the
// maxima are signed, as the reference's reducer takes them (max e over max
// exact, from the max identity -DBL_MAX).
int lpm_err_norms_scalar(const double* err, const double* exact, const double* wt, int64_t n, double* out) {
  if (n < 1 || !err || !exact || !wt || !out) return fail(LPM_ERR_ARG, "lpm_err_norms_scalar: bad arguments");
  double l1n = 0, l1d = 0, l2n = 0, l2d = 0, lin = -DBL_MAX, lid = -DBL_MAX;
  for (int64_t i = 0; i < n; ++i) {
    const double e = err[i], x = exact[i];
    l1n += std::fabs(e) * wt[i];
    l1d += std::fabs(x) * wt[i];
    l2n += (e * e) * wt[i];
    l2d += (x * x) * wt[i];
    lin = e > lin ? e : lin;
    lid = x > lid ? x : lid;
  }
  out[0] = l1n / l1d;
  out[1] = std::sqrt(l2n / l2d);
  out[2] = lin / lid;
  return ok();
}

// SphereVelocityTangentTestFunctor (src/LpmSphereTestKernels.hpp:11-25):
// udotx[i] = x_i . u_i (SphereGeometry::dot, left to right).
int lpm_velocity_tangent(const double* x, const double* u, int64_t n, double* udotx) {
  if (n < 0 || (n && (!x || !u || !udotx))) return fail(LPM_ERR_ARG, "lpm_velocity_tangent: bad arguments");
  for (int64_t i = 0; i < n; ++i)
    udotx[i] = x[3 * i] * u[3 * i] + x[3 * i + 1] * u[3 * i + 1] + x[3 * i + 2] * u[3 * i + 2];
  return ok();
}

}  // extern "C"
