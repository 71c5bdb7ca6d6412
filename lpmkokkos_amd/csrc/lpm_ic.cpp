// Initial conditions (host).  Compiled with -ffp-contract=off.
#include "lpm_ic.hpp"

#include <cmath>

namespace lpm {

namespace {
uint64_t splitmix64_next(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
}  // namespace

Xoshiro256ss::Xoshiro256ss(uint64_t seed) {
  uint64_t x = seed;
  for (auto& w : s_) w = splitmix64_next(x);
}

uint64_t Xoshiro256ss::next() {
  const uint64_t out = rotl64(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl64(s_[3], 45);
  return out;
}

double Xoshiro256ss::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

double Xoshiro256ss::normal() {
  const double a = 1.0 - uniform();  // (0, 1]
  const double b = uniform();
  return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
}

double CompositeVorticity::eval(const double& x, const double& y, const double& z) const {
  const Vec3 p{{x, y, z}};
  double zeta = 0.0;
  if (sbr != 0.0) zeta += sbr * (2 * (2 * kPi) * z);
  if (rh_n > 0) {
    const double lat = sphere::latitude(p), lon = sphere::longitude(p);
    const double cl = std::cos(lat), sl = std::sin(lat);
    zeta += 2 * rh_w * sl - rh_K * (rh_n + 1) * (rh_n + 2) * std::pow(cl, rh_n) * sl * std::cos(rh_n * lon);
  }
  const double s2 = sigma * sigma;
  for (size_t k = 0; k < amps.size(); ++k) zeta += amps[k] * std::exp(-(1.0 - sphere::dot(p, centers[k])) / s2);
  return zeta;
}

void CompositeVorticity::draw_vortices(uint64_t seed, int nvort, double amp_scale, double sigma_) {
  Xoshiro256ss rng(seed);
  centers.clear();
  amps.clear();
  for (int k = 0; k < nvort; ++k) {
    Vec3 c;
    c[0] = rng.normal();
    c[1] = rng.normal();
    c[2] = rng.normal();
    sphere::normalize(c);
    centers.push_back(c);
    amps.push_back(amp_scale * rng.normal());
  }
  sigma = sigma_;
}

double sph_harm54(const Vec3& x) {
  const double lon = sphere::longitude(x);
  const double z = x[2];
  const double p54 = z * (z * z - 1.0) * (z * z - 1.0);
  return 30 * std::cos(4 * lon) * p54;
}

Vec3 rh54_velocity(const Vec3& x) {
  const double lat = sphere::latitude(x), lon = sphere::longitude(x);
  const double cl = std::cos(lat), sl = std::sin(lat), cn = std::cos(lon), sn = std::sin(lon);
  const double cl3 = cl * cl * cl;
  // zonal and meridional components, then the Cartesian velocity
  const double us = 0.5 * std::cos(4 * lon) * cl3 * (5 * std::cos(2 * lat) - 3);
  const double vs = 4 * cl3 * sl * std::sin(4 * lon);
  return Vec3{{-us * sn - vs * sl * cn, us * cn - vs * sl * sn, vs * cl}};
}

}  // namespace lpm
