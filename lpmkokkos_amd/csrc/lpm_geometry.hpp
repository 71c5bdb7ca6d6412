// Sphere geometry used by the host-side mesh build and initial conditions.
//
// The operations and their evaluation order follow SphereGeometry in
// src/LpmGeometry.hpp:203-469 and atan4 in src/LpmUtilities.hpp:14-46, so a
// mesh built here is bit-identical to the reference's (this file is compiled
// with -ffp-contract=off).
#pragma once

#include <cmath>

namespace lpm {

constexpr double kPi = 3.1415926535897932384626433832795027975;  // src/LpmDefs.hpp:82

struct Vec3 {
  double v[3];
  double& operator[](int i) { return v[i]; }
  const double& operator[](int i) const { return v[i]; }
};

namespace sphere {

inline double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

inline Vec3 cross(const Vec3& a, const Vec3& b) {
  return Vec3{{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}

inline double mag(const Vec3& a) { return std::sqrt(dot(a, a)); }

// Scale by the reciprocal of the magnitude (SphereGeometry::normalize).
inline void normalize(Vec3& a) {
  const double r = 1.0 / mag(a);
  for (int i = 0; i < 3; ++i) a[i] *= r;
}

// Great-circle distance as atan2(|a x b|, a.b).
inline double distance(const Vec3& a, const Vec3& b) { return std::atan2(mag(cross(a, b)), dot(a, b)); }

// Normalised half-sum of two points.
inline Vec3 midpoint(const Vec3& a, const Vec3& b) {
  Vec3 m{{0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (a[2] + b[2])}};
  normalize(m);
  return m;
}

// Running sum of the corners, scaled by 1.0/n, then normalised.
inline Vec3 barycenter(const Vec3* pts, int n) {
  Vec3 c{{0.0, 0.0, 0.0}};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) c[k] += pts[i][k];
  const double s = 1.0 / n;
  for (int k = 0; k < 3; ++k) c[k] *= s;
  normalize(c);
  return c;
}

// Spherical triangle area by L'Huilier's theorem.
inline double tri_area(const Vec3& a, const Vec3& b, const Vec3& c) {
  const double s1 = distance(a, b), s2 = distance(b, c), s3 = distance(c, a);
  const double hp = 0.5 * (s1 + s2 + s3);
  const double zz = std::tan(0.5 * hp) * std::tan(0.5 * (hp - s1)) * std::tan(0.5 * (hp - s2)) *
                    std::tan(0.5 * (hp - s3));
  return 4 * std::atan(std::sqrt(zz));
}

// Fan of triangles (centre, corner i, corner i+1), accumulated in corner order.
inline double polygon_area(const Vec3& ctr, const Vec3* corners, int n) {
  double total = 0;
  for (int i = 0; i < n; ++i) total += tri_area(ctr, corners[i], corners[(i + 1) % n]);
  return total;
}

// atan2 folded into [0, 2 pi), with the reference's handling of the axes.
inline double atan4(double y, double x) {
  if (x == 0.0) {
    if (y > 0.0) return 0.5 * kPi;
    if (y < 0.0) return 1.5 * kPi;
    return 0.0;
  }
  if (y == 0) return x > 0.0 ? 0.0 : (x < 0.0 ? kPi : 0.0);
  const double th = std::atan2(std::fabs(y), std::fabs(x));
  if (x > 0.0 && y > 0.0) return th;
  if (x < 0.0 && y > 0.0) return kPi - th;
  if (x < 0.0 && y < 0.0) return kPi + th;
  return 2.0 * kPi - th;
}

inline double latitude(const Vec3& p) { return std::atan2(p[2], std::sqrt(p[0] * p[0] + p[1] * p[1])); }
inline double longitude(const Vec3& p) { return atan4(p[1], p[0]); }

}  // namespace sphere
}  // namespace lpm
