// Vorticity initial conditions behind the reference's plugin interface
// VorticityInitialCondition::eval(x, y, z) (src/LpmVorticityGallery.hpp:13-40).
// SolidBodyRotation is the reference's; the Rossby-Haurwitz wave and the
// Gaussian-vortex sum are new fields that plug into the same interface.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "lpm_geometry.hpp"

namespace lpm {

struct VorticityInitialCondition {
  typedef std::shared_ptr<VorticityInitialCondition> ptr;
  virtual ~VorticityInitialCondition() {}
  virtual double eval(const double& x, const double& y, const double& z) const = 0;
  virtual std::string name() const { return std::string(); }
};

// Counter-based generator: SplitMix64 (Steele, Lea and Flood 2014) expands a
// 64-bit seed into the state of xoshiro256** (Blackman and Vigna 2018).
class Xoshiro256ss {
 public:
  explicit Xoshiro256ss(uint64_t seed);
  uint64_t next();
  double uniform();  // [0, 1), 53 random bits
  double normal();   // Box-Muller, cosine branch, two uniforms per value
 private:
  uint64_t s_[4];
};

// Sum of the three components, added in this order starting from 0:
// solid-body rotation 2*(2 pi)*z scaled by `sbr`, a Rossby-Haurwitz wave of
// wavenumber rh_n (skipped when rh_n <= 0), then each Gaussian vortex
// amp_k * exp(-(1 - x.c_k) / sigma^2) in index order.
class CompositeVorticity : public VorticityInitialCondition {
 public:
  double sbr = 0.0;
  int rh_n = 0;
  double rh_w = 0.0, rh_K = 0.0;
  std::vector<Vec3> centers;
  std::vector<double> amps;
  double sigma = 1.0;

  double eval(const double& x, const double& y, const double& z) const override;
  std::string name() const override { return "composite"; }

  // Draw `nvort` vortices from `seed`: per vortex three normals normalised
  // to the sphere, then amp_scale times one normal.
  void draw_vortices(uint64_t seed, int nvort, double amp_scale, double sigma_);
};

// SolidBodyRotation (src/LpmVorticityGallery.hpp:25-40): zeta = 2 * (2 pi) * z.
struct SolidBodyRotation : public CompositeVorticity {
  SolidBodyRotation() { sbr = 1.0; }
  std::string name() const override { return "rotation"; }
};

// Rossby-Haurwitz wave of wavenumber n with parameters w and K.
struct RossbyHaurwitzWave : public CompositeVorticity {
  RossbyHaurwitzWave(int n, double w, double K) {
    rh_n = n;
    rh_w = w;
    rh_K = K;
  }
  std::string name() const override { return "rossby_haurwitz"; }
};

// The Poisson problem of SpherePoisson (src/LpmSpherePoisson.hpp:60-80,
// src/LpmRossbyWaves.hpp:18-53): right-hand side f = Y_5^4 =
// 30 cos(4 lon) P_5^4(z) with P_5^4(z) = z (z^2 - 1)^2, exact stream function
// f / 30, and the velocity it induces (RH54Velocity).
double sph_harm54(const Vec3& x);
Vec3 rh54_velocity(const Vec3& x);

}  // namespace lpm
