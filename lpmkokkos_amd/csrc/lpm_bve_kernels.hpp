// Device kernels of the spherical Biot-Savart direct sum and the RK4 stage
// arithmetic, for gfx950 (CDNA4, 64-lane wavefronts).
//
// Pair sum (src/LpmBVEKernels.hpp:21-44, 49-102, 188-234):
//   u_i   = sum_j (x_i x y_j) q_j / (1 - x_i.y_j),  q_j = -zeta_j A_j / (4 pi)
//   psi_i = sum_j q_j log(1 - x_i.y_j)
// over the active (unmasked) faces j, with j = i skipped for face targets.
//
// Layout and work split:
//  * Targets are SoA FP64 (x, y, z) arrays; each lane owns R targets, so a
//    256-lane workgroup covers 256*R consecutive targets and keeps them in
//    registers for the whole source sweep.
//  * Sources are compacted active faces in the tree order (which is spatial:
//    consecutive 4^k leaves are one level-(L-k) triangle), padded to whole
//    tiles of kTile.  By default a workgroup stages its source tiles through
//    LDS, double-buffered: every lane copies one record of tile t+1 (one
//    coalesced load) while the waves sum tile t, with a barrier per tile;
//    waves read the records with uniform-address (broadcast) LDS loads.
//    The alternative staging reads each record with a uniform address
//    straight from memory, which the compiler issues as scalar loads.
//  * The source range is cut into `splits` contiguous tile ranges
//    (grid.y); each split writes FP64 partial sums that a finalize kernel
//    adds in split order, so results do not depend on scheduling.
//
// Arithmetic forms:
//  * FORM_DOT  : FP64, literal 1 - x.y (the reference's denominator).
//  * FORM_CHORD: FP64, 1 - x.y = |y - x|^2 / 2 (exact on the sphere, no
//    cancellation); u = x x sum_j s_j (y_j - x_i).
//  * FP32C     : FP32 pair terms in tile-local coordinates.  Each tile has an
//    FP64 centre c; a source is stored as the FP32 offset y - c, a target
//    forms D = fl32(c - x) once per tile in FP64, and d = D + (y - c) is
//    accurate to ~6e-8 of the tile radius instead of ~6e-8 of |x|.  Tile
//    partials are accumulated in FP32 and combined in FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace lpm {
namespace dev {

constexpr int kBlock = 256;  // lanes per workgroup (4 wavefronts)
constexpr int kTile = 256;   // sources per tile
constexpr int kNC = 4;       // partial components: Sx, Sy, Sz, P

enum Form { FORM_DOT = 0, FORM_CHORD = 1 };

// Strided view of an [n][3] or SoA position array: element (i, c) is at
// p[i * is + c * cs].
struct Pos3 {
  const double* p;
  long is, cs;
  __device__ __forceinline__ double at(long i, int c) const { return p[i * is + c * cs]; }
};
struct MutPos3 {
  double* p;
  long is, cs;
  __device__ __forceinline__ double& at(long i, int c) const { return p[i * is + c * cs]; }
};

}  // namespace dev
}  // namespace lpm
