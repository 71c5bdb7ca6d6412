"""MI355X-native BVE Biot-Savart direct sum (the hot path of pbosler/lpmKokkos).

The compute lives in liblpm_bve.so (hand-written HIP for gfx950 behind the C
ABI of include/lpm_bve.h); this package is its Python front end.
"""
from ._lib import (PREC_FP32C, PREC_FP64, SCHEME_LAGR, SCHEME_REF, SEED_CUBED_SPHERE, SEED_ICOS_TRI,  # noqa: F401
                   VTK_LEGACY_ASCII, VTK_LEGACY_BINARY, VTK_XML_ASCII, VTK_XML_BINARY, LpmError, lib)
from .bve import (BVERK4, BVESphere, Context, CubedSphereSeed, ErrNorms, IcosTriSphereSeed, MeshSeed,  # noqa: F401
                  RossbyHaurwitzWave, SolidBodyRotation, SpherePoisson, TreeMesh, VorticityInitialCondition,
                  area_mean, gauss_params, poisson_y54, velocity_tangent, write_mesh_nc, write_mesh_vtk)
