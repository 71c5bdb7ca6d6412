/*
 * lpm_bve.h -- C ABI of the MI355X-native BVE Biot-Savart library
 * (liblpm_bve.so).
 *
 * Plain pointers and sizes only.  Every call returns an lpm_status; on a
 * non-zero status lpm_last_error() (thread-local) describes the failure.
 * No exception crosses this boundary.
 *
 * Layouts: positions and velocities are [n][3] row-major doubles (the
 * reference's host layout of View<double*[3]>, src/LpmDefs.hpp:43-47);
 * scalars are [n] doubles; masks are [n] bytes (non-zero = divided face,
 * excluded as a source, src/LpmFaces.cpp:244-247).  "Tree faces" means all
 * faces of the refinement tree, parents included (src/LpmBVERK4.cpp:21-36).
 */
#ifndef LPM_BVE_H
#define LPM_BVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LPM_OK = 0,
  LPM_ERR_ARG = 1,     /* invalid argument or size */
  LPM_ERR_HIP = 2,     /* HIP runtime error */
  LPM_ERR_NOMEM = 3,   /* device allocation failed */
  LPM_ERR_STATE = 4,   /* call not valid in the current state */
  LPM_ERR_BUDGET = 5,  /* request exceeds the admission memory budget */
  LPM_ERR_IO = 6       /* file could not be read */
} lpm_status;

/* Arithmetic of the pair sums. */
typedef enum {
  LPM_PREC_FP64 = 0,  /* FP64 throughout */
  LPM_PREC_FP32C = 1  /* FP32 pair terms from tile-local offsets, FP64 tile partials */
} lpm_prec;

/* Time-stepping scheme. */
typedef enum {
  /* The reference step as written (src/LpmBVERK4_Impl.hpp:42-130): kernel
   * denominator 1 - x.y, vorticity integrated with dzeta = -2 Omega w dt, the
   * face-vorticity combination (1,2,4,4) of :120-121, no renormalisation. */
  LPM_SCHEME_REF = 0,
  /* Projected RK4: stage and updated positions normalised to the unit
   * sphere, zeta = (zeta0 + 2 Omega z0) - 2 Omega z, denominator |y-x|^2/2. */
  LPM_SCHEME_LAGR = 1
} lpm_scheme;

/* VTK polydata layouts of lpm_bve_write_vtk. */
typedef enum {
  LPM_VTK_LEGACY_ASCII = 0,  /* BVESphere::outputVtk (src/LpmBVESphere.cpp:90-113): legacy .vtk */
  LPM_VTK_LEGACY_BINARY = 1,
  LPM_VTK_XML_ASCII = 2,     /* Polymesh2dVtkInterface::write + addFieldsToVtk: .vtp */
  LPM_VTK_XML_BINARY = 3     /* .vtp, inline base64 */
} lpm_vtk_format;

/* Status text, version and device count: no counterpart in the reference
 * (whose errors are C++ exceptions, LPM_THROW_IF in src/LpmDefs.hpp); for
 * callers across the C ABI. */
const char* lpm_last_error(void);
int lpm_version(void); /* major*10000 + minor*100 + patch */
int lpm_device_count(int* count);

/* ------------------------------------------------------------------ */
/* Mesh build (host).                                                  */
/* ------------------------------------------------------------------ */

/* The spherical mesh seeds (src/LpmMeshSeed.hpp:111-147): the icosahedral
 * triangle seed and the cubed-sphere quadrilateral seed (4 vertices per
 * face; a divided face's centre becomes its children's shared corner,
 * src/LpmFaces.cpp:251-381). */
typedef enum {
  LPM_SEED_ICOS_TRI = 0,     /* IcosTriSphereSeed, mesh_seeds/icosTriSphereSeed.dat */
  LPM_SEED_CUBED_SPHERE = 1  /* CubedSphereSeed, mesh_seeds/cubedSphereSeed.dat */
} lpm_seed_kind;

/* MeshSeed<Seed>::setMaxAllocations (src/LpmMeshSeed.cpp:296-305, counts of
 * :369-391): vertex, edge and tree-face counts of a level-`level` tree. */
int lpm_seed_alloc(int seed_kind, int level, int64_t* nverts, int64_t* nedges, int64_t* nfaces);
/* The same for the icosahedral seed. */
int lpm_icos_alloc(int level, int64_t* nverts, int64_t* nedges, int64_t* nfaces);

/* PolyMesh2d<IcosTriSphereSeed>::treeInit (src/LpmPolyMesh2d.cpp:30-44) with
 * the built-in seed (seed_path == NULL) or a seed file in the reference's
 * format (src/LpmMeshSeed.cpp:20-232).  Arrays are sized by lpm_icos_alloc:
 * vx, vlag [nverts][3]; fx, flag [nfaces][3]; farea, fmask [nfaces];
 * fverts [nfaces][3]; fparent [nfaces].  Any output may be NULL.
 * counts (non-NULL) receives {nverts, nedges, nfaces, nleaves}. */
int lpm_icos_tree_init(int level, const char* seed_path, double* vx, double* vlag, double* fx, double* flag,
                       double* farea, uint8_t* fmask, int32_t* fverts, int32_t* fparent, int64_t* counts);
/* PolyMesh2d<Seed>::treeInit for either seed (the seed file, if given, in the
 * reference's format for that seed); fverts is [nfaces][3] for the
 * icosahedral seed and [nfaces][4] for the cubed sphere. */
int lpm_seed_tree_init(int seed_kind, int level, const char* seed_path, double* vx, double* vlag, double* fx,
                       double* flag, double* farea, uint8_t* fmask, int32_t* fverts, int32_t* fparent,
                       int64_t* counts);

/* NcWriter::writePolymesh (src/LpmNetCDF_Impl.hpp:46-218) of the treeInit mesh
 * (as tests/LpmNetCDFTest.cpp:95-104 writes icostri_sphere.nc): dimensions
 * ndim, nverts, nedges, nfaces, nfaceverts, two, four; coordinates, edge and
 * face topology and tree, areas and mask; attributes MeshSeed, FaceKind, Geom,
 * baseTreeDepth.  A netCDF classic (64-bit offset) file. */
int lpm_mesh_write_nc(int level, const char* seed_path, const char* path);
/* The same for either seed (nfaceverts 4, MeshSeed "CubedSphereSeed",
 * FaceKind "quad" for the cubed sphere). */
int lpm_seed_mesh_write_nc(int seed_kind, int level, const char* seed_path, const char* path);

/* PolyMesh2d::outputVtk (src/LpmPolyMesh2d.cpp:79-85) of the treeInit mesh:
 * vertices, leaf faces and their areas ("area" cell data), in one of the
 * lpm_vtk_format layouts. */
int lpm_mesh_write_vtk(int level, const char* seed_path, const char* path, int format);
int lpm_seed_mesh_write_vtk(int seed_kind, int level, const char* seed_path, const char* path, int format);

/* PolyMeshReader (src/LpmNetCDF.cpp:14-240): the particle arrays of a mesh
 * file written by lpm_mesh_write_nc or lpm_bve_checkpoint, in the layout of
 * lpm_seed_tree_init.  Call with NULL arrays to get counts {nverts, nedges,
 * nfaces, nleaves} (nedges -1 when the file holds no edges); fverts and
 * fparent need a file with topology; fverts receives the file's nfaceverts
 * per face (4 when nverts = nleaves + 2, the cubed sphere; else 3). */
int lpm_mesh_read_nc(const char* path, double* vx, double* vlag, double* fx, double* flag, double* farea,
                     uint8_t* fmask, int32_t* fverts, int32_t* fparent, int64_t* counts);

/* ------------------------------------------------------------------ */
/* Vorticity initial conditions (host), the                           */
/* VorticityInitialCondition::eval(x,y,z) interface of                 */
/* src/LpmVorticityGallery.hpp:13-40, as one composite field:          */
/*   zeta = sbr * 2 * (2 pi) * z                       (SolidBodyRotation)  */
/*        + 2 w sin(lat) - K (n+1)(n+2) cos^n(lat) sin(lat) cos(n lon) (RH, n>0) */
/*        + sum_k amp_k exp(-(1 - x.c_k) / sigma^2)    (Gaussian vortices)  */
/* ------------------------------------------------------------------ */

/* Seeded vortex parameters: SplitMix64 -> xoshiro256**, Box-Muller normals;
 * for each k: three normals normalised to the sphere (centre), then
 * amp_scale times one normal (amplitude). centers [nvort][3], amps [nvort]. */
int lpm_ic_gauss_params(uint64_t seed, int nvort, double amp_scale, double* centers, double* amps);

/* VorticityInitialCondition::eval (src/LpmVorticityGallery.hpp:13-40) of the
 * composite field above at n points xyz [n][3] -> zeta [n]. */
int lpm_ic_eval(const double* xyz, int64_t n, double* zeta, double sbr, int rh_n, double rh_w, double rh_K,
                int nvort, const double* centers, const double* amps, double sigma);

/* The same generator's raw draws: nuniform values in [0,1), then nnormal
 * standard normals (used to draw ensemble and request parameters).  No
 * counterpart in the reference; for the seeded batching and serving
 * workloads. */
int lpm_rng_draws(uint64_t seed, int64_t nuniform, double* uniform, int64_t nnormal, double* normal);

/* The Poisson problem of SpherePoisson (src/LpmSpherePoisson.hpp:60-80):
 * f = Y_5^4(x) = 30 cos(4 lon) z (z^2-1)^2 (exact stream function f/30) and
 * its velocity RH54Velocity (src/LpmRossbyWaves.hpp:39-53) at n points.
 * f [n] and u [n][3] may each be NULL. */
int lpm_poisson_y54(const double* xyz, int64_t n, double* f, double* u);

/* Area-weighted mean of zeta over unmasked faces, summed in index order (the
 * mean that the Gaussian-vortex initial fields subtract; no separate function
 * in the reference). */
int lpm_area_mean(const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, double* mean);

/* ------------------------------------------------------------------ */
/* Diagnostics (host).                                                 */
/* ------------------------------------------------------------------ */

/* ErrNorms(err, appx, exact, wt) for vector fields (src/LpmErrorNorms.cpp:15-45):
 * err = appx - exact ([n][3]; err may be NULL), then, in index order,
 * out = {l1, l2, linf}: l1 = sum|e|w / sum|x|w, l2 = sqrt(sum|e|^2 w / sum|x|^2 w),
 * linf = max|e| / max|x|. */
int lpm_err_norms_vec(double* err, const double* appx, const double* exact, const double* wt, int64_t n,
                      double* out);
/* ErrNorms(err, exact, wt) for a scalar error field err = appx - exact
 * (src/LpmErrorNorms.cpp:56-70): l1 and l2 with |e| and |x|; linf = max e /
 * max exact with SIGNED maxima, as the reference's reducer takes them. */
int lpm_err_norms_scalar(const double* err, const double* exact, const double* wt, int64_t n, double* out);
/* SphereVelocityTangentTestFunctor (src/LpmSphereTestKernels.hpp:11-25):
 * udotx[i] = x_i . u_i for x, u [n][3]. */
int lpm_velocity_tangent(const double* x, const double* u, int64_t n, double* udotx);

/* ------------------------------------------------------------------ */
/* Device functors (device pointers, asynchronous on `stream`, NULL =  */
/* the null stream: a call queues its kernels and returns).  FP64,     */
/* denominator 1 - x.y, as the reference.  Scratch device memory comes */
/* from a per-device workspace that grows on demand (growing waits for */
/* the previous call and allocates); each call waits on the device for */
/* the previous call's kernels before reusing it, on any stream.       */
/* ------------------------------------------------------------------ */

/* BVEVertexVelocity (src/LpmBVEKernels.hpp:163-186): u[nv][3] at targets
 * vx[nv][3] from the unmasked faces among fx[nf][3]. */
int lpm_bve_vertex_velocity(double* u, const double* vx, int64_t nv, const double* fx, const double* fzeta,
                            const double* farea, const uint8_t* fmask, int64_t nf, void* stream);
/* BVEFaceVelocity (src/LpmBVEKernels.hpp:285-305): u[nf][3] at every face,
 * the face itself excluded as a source. */
int lpm_bve_face_velocity(double* u, const double* fx, const double* fzeta, const double* farea,
                          const uint8_t* fmask, int64_t nf, void* stream);
/* BVEVertexStreamFn / BVEFaceStreamFn (src/LpmBVEKernels.hpp:140-161, 263-283). */
int lpm_bve_vertex_streamfn(double* psi, const double* vx, int64_t nv, const double* fx, const double* fzeta,
                            const double* farea, const uint8_t* fmask, int64_t nf, void* stream);
int lpm_bve_face_streamfn(double* psi, const double* fx, const double* fzeta, const double* farea,
                          const uint8_t* fmask, int64_t nf, void* stream);
/* BVEVertexSolve / BVEFaceSolve (src/LpmBVEKernels.hpp:110-138, 236-261): the
 * stream function and the velocity together, in one pass over the sources
 * (the reference runs the two reductions one after the other); argument
 * order as the reference's constructors: psi, u, targets, sources, ... */
int lpm_bve_vertex_solve(double* psi, double* u, const double* vx, int64_t nv, const double* fx,
                         const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, void* stream);
int lpm_bve_face_solve(double* psi, double* u, const double* fx, const double* fzeta, const double* farea,
                       const uint8_t* fmask, int64_t nf, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The functor evaluation above (psi and/or u, either may
 * be NULL but not both; collocated != 0: the targets are the nf faces fx
 * themselves, each face's own source skipped, as lpm_bve_face_solve; else
 * nt targets tx, as lpm_bve_vertex_solve), with the same kernels and launch
 * shape, so the results are bit-identical to those entry points'.  Its
 * scratch (source records, split partials, source map) lives in a fixed
 * static device array of LPM_FUNCTOR_STATIC_SCRATCH_BYTES that each device
 * holds until the process ends; it calls no hipMalloc or hipFree and never
 * waits.  LPM_ERR_ARG when the sizes need more scratch than that (the
 * evaluations of trees up to level 6, 109,220 faces, fit).  Every call on a
 * device shares that device's array: calls that overlap on one device race
 * (queue them on one stream). */
#define LPM_FUNCTOR_STATIC_SCRATCH_BYTES (64ull << 20)
int lpm_bve_functor_static(double* psi, double* u, const double* tx, int64_t nt, const double* fx,
                           const double* fzeta, const double* farea, const uint8_t* fmask, int64_t nf, int collocated,
                           void* stream);

/* ------------------------------------------------------------------ */
/* Solver context: the BVESphere fields and BVERK4 stage buffers of    */
/* one problem, resident in device memory (src/LpmBVESphere.hpp:18-80, */
/* src/LpmBVERK4.hpp:13-65).  One host thread uses a context at a      */
/* time; contexts may run concurrently, each on its own stream.        */
/* ------------------------------------------------------------------ */
typedef struct lpm_bve_ctx lpm_bve_ctx;

/* The BVESphere(nmaxverts, nmaxedges, nmaxfaces) and BVERK4(dt, Omega)
 * constructors (src/LpmBVESphere.cpp:10-49, src/LpmBVERK4.cpp:5-37) and their
 * destructors: a context with its own stream on `device`; sizes come with the
 * state. */
int lpm_bve_create(int device, lpm_bve_ctx** ctx);
/* A context that queues its work on a caller's stream (a hipStream_t of
 * `device`, not owned: it must outlive the context).  Contexts sharing a
 * stream, e.g. the members of one ensemble on the ensemble's stream
 * (lpm_ens_stream), are ordered by the stream and need no cross-stream
 * events. */
int lpm_bve_create_on_stream(int device, void* stream, lpm_bve_ctx** ctx);
int lpm_bve_destroy(lpm_bve_ctx* ctx);

/* Upload a particle state (host arrays; copied): the physVerts, relVortVerts,
 * physFaces, relVortFaces, faces.area and faces.mask views that treeInit and
 * init_vorticity fill (src/LpmBVESphere.cpp:126-149) and advance_timestep
 * reads (src/LpmBVERK4.hpp:29-34).  vx [nv][3], vzeta [nv]; fx [nf][3],
 * fzeta, farea, fmask [nf] over tree faces.  Velocities become invalid until
 * lpm_bve_solve, lpm_bve_set_velocity or the first step. */
int lpm_bve_set_state(lpm_bve_ctx* ctx, int64_t nv, const double* vx, const double* vzeta, int64_t nf,
                      const double* fx, const double* fzeta, const double* farea, const uint8_t* fmask);

/* Vorticity field evaluated on the device: the composite of lpm_ic_eval
 * (solid-body rotation scaled by sbr, Rossby-Haurwitz wave rh_n > 0,
 * nvort Gaussian vortices with host arrays centers [nvort][3], amps [nvort]). */
typedef struct lpm_ic_params {
  double sbr;
  int rh_n;
  double rh_w, rh_K;
  int nvort;
  const double* centers;
  const double* amps;
  double sigma;
  int subtract_mean; /* subtract the area-weighted mean over active faces */
} lpm_ic_params;

/* The state of MeshSeed<IcosTriSphereSeed> + treeInit(level) +
 * init_vorticity(ic) (src/LpmPolyMesh2d.cpp:30-44, src/LpmBVESphere.cpp:126-149)
 * without host arrays: the level's mesh comes from a per-device cache in
 * device memory (built by lpm_icos_tree_init's code and uploaded on first use,
 * shared read-only by every context on the device) and the field is
 * evaluated on the device, in FP64 (values agree with lpm_ic_eval's to
 * rounding, not bit for bit).  Positions, numbering and areas are those of
 * lpm_icos_tree_init(level, NULL, ...); nf counts every tree face. */
int lpm_bve_set_state_level(lpm_bve_ctx* ctx, int level, const lpm_ic_params* ic);
/* The same for MeshSeed<SeedType> of either seed (lpm_seed_kind): the
 * cubed-sphere tree of CubedSphereSeed (src/LpmMeshSeed.hpp:111) with
 * lpm_seed_tree_init's positions, numbering and areas, cached per (device,
 * seed, level).  lpm_bve_set_state_level(ctx, level, ic) is this call with
 * LPM_SEED_ICOS_TRI. */
int lpm_bve_set_state_seed_level(lpm_bve_ctx* ctx, int seed_kind, int level, const lpm_ic_params* ic);

/* Free the cached meshes of `device` (all devices if device < 0) that no
 * context uses; *kept (may be NULL) receives the number still in use.  No
 * counterpart in the reference (it rebuilds the mesh per object). */
int lpm_mesh_cache_clear(int device, int* kept);

/* Provide the entry velocity: the velocityVerts / velocityFaces views that
 * advance_timestep requires to hold the velocity at the current state
 * (src/LpmBVERK4.hpp:32-33, stage 1 at src/LpmBVERK4_Impl.hpp:61-68).
 * vu [nv][3], fu [nf][3]. */
int lpm_bve_set_velocity(lpm_bve_ctx* ctx, const double* vu, const double* fu);

/* Velocity (and, if want_psi, stream function) at the current state, as the
 * solves of BVESphere::init_vorticity (src/LpmBVESphere.cpp:151-168).
 * scheme selects the kernel denominator (REF: 1 - x.y; LAGR: |y-x|^2/2). */
int lpm_bve_solve(lpm_bve_ctx* ctx, int prec, int scheme, int want_psi);
/* Stream function alone at the current state: the reference driver's
 * post-timestep BVEVertexStreamFn / BVEFaceStreamFn (tests/LpmBVETest.cpp:155-159,
 * src/LpmBVEKernels.hpp:140-161, 263-283).  The velocity the context holds
 * (the next step's entry velocity) is left as it is, so steps after it are
 * the steps without it, bit for bit. */
int lpm_bve_streamfn(lpm_bve_ctx* ctx, int prec, int scheme);

/* `nsteps` calls of BVERK4::advance_timestep (src/LpmBVERK4_Impl.hpp:42-130)
 * on the resident state.  On return the velocity is valid at the new state.
 * LPM_PREC_FP32C requires LPM_SCHEME_LAGR.  Steps are replayed from a
 * captured hipGraph when use_graph != 0. */
int lpm_bve_advance(lpm_bve_ctx* ctx, double dt, double omega, int scheme, int prec, int nsteps, int use_graph);

/* Copy the resident state to host arrays (any pointer may be NULL), as
 * BVESphere::updateHost does for its host mirrors (src/LpmBVESphere.cpp:77-88).
 * Blocks until the context's stream has drained. */
int lpm_bve_get_state(lpm_bve_ctx* ctx, double* vx, double* vzeta, double* vu, double* vpsi, double* fx,
                      double* fzeta, double* fu, double* fpsi);

/* Checkpoint: the state in a netCDF file that lpm_bve_restore resumes from
 * bit for bit.  The mesh variables of NcWriter::writePolymesh (all of them
 * when the tree is known: `level` >= 0 names it, -1 takes a level-built
 * state's own; else positions, areas and mask only), the BVESphere fields
 * under their view labels (relVortVerts, absVortVerts, streamFnVerts,
 * velocityVerts, and *Faces; src/LpmBVESphere.cpp:13-20), and attributes
 * time, Omega, velocity_prec, velocity_form, streamfn_valid, absvort_valid,
 * absvort_omega.  The absolute vorticity is that of the first step at rate
 * `omega` (computed now if no step has run at that rate). */
int lpm_bve_checkpoint(lpm_bve_ctx* ctx, const char* path, int level, double time, double omega);

/* Resume from a checkpoint: the state, its velocity (with the precision and
 * form that produced it), psi and absolute vorticity, so that the next steps
 * are those the checkpointed context would have taken.  time and omega
 * (non-NULL) receive the stored values.  No counterpart in the reference,
 * whose restart (PolyMesh2d(const PolyMeshReader&), src/LpmPolyMesh2d.cpp:20-28)
 * restores the mesh only; for callers that stop and resume runs (the
 * serving loop's drain and resume). */
int lpm_bve_restore(lpm_bve_ctx* ctx, const char* path, double* time, double* omega);

/* VTK polydata of the state on the leaf faces of its tree (`level` as for
 * lpm_bve_checkpoint; required unless the state was built by level).  Point
 * data relVort, absVort, streamFn, velocity; cell data area, relVort, absVort,
 * streamFn, velocity (legacy names; the XML layouts use addFieldsToVtk's
 * relvort, absvort, stream_fn, velocity).  streamFn is zero unless the last
 * solve computed psi. */
int lpm_bve_write_vtk(lpm_bve_ctx* ctx, const char* path, int level, double omega, int format);

/* A caller's array for lpm_bve_write_vtk_arrays: at vertices [nverts][ncomp]
 * or at faces [nfaces][ncomp] in tree order (leaf rows are written). */
typedef struct {
  const char* name; /* 1-255 bytes; written as VTK encodes names (%XX in legacy files) */
  int ncomp;        /* 1..9 */
  int on_faces;     /* 0: point data, 1: cell data */
  const double* data;
} lpm_vtk_array;

/* lpm_bve_write_vtk with `narrays` more arrays after the state's own, as the
 * reference's BVE driver adds its error fields with addVectorPointData /
 * addVectorCellData (tests/LpmBVETest.cpp:98-106, 228-238). */
int lpm_bve_write_vtk_arrays(lpm_bve_ctx* ctx, const char* path, int level, double omega, int format,
                             const lpm_vtk_array* arrays, int narrays);

/* lpm_seed_mesh_write_vtk with `narrays` arrays after the leaf areas, host
 * only: SpherePoisson::outputVtk (src/LpmSpherePoisson.hpp:510-531), which
 * adds the solution, exact solution and error fields to the mesh. */
int lpm_seed_mesh_write_vtk_arrays(int seed_kind, int level, const char* seed_path, const char* path, int format,
                                   const lpm_vtk_array* arrays, int narrays);

/* Total circulation sum over unmasked faces of zeta * area (FP64, device):
 * the conservation check of the Lagrangian scheme; the reference computes it
 * in its drivers from the host mirrors. */
int lpm_bve_circulation(lpm_bve_ctx* ctx, double* circ);

/* Context control, launch shape and timing: no counterpart in the reference
 * (Kokkos fences and chooses team sizes itself); for asynchronous callers,
 * the serving loop and measurement.  Block until all work queued on the
 * context has finished. */
int lpm_bve_synchronize(lpm_bve_ctx* ctx);
/* Non-blocking: *done = 1 when the context's asynchronous work (solves, steps,
 * ensemble steps) queued so far has finished. */
int lpm_bve_query(lpm_bve_ctx* ctx, int* done);

/* Sizes of the resident problem: {nv, nf, nsources, ntargets}. */
int lpm_bve_sizes(lpm_bve_ctx* ctx, int64_t* sizes4);

/* Kernel launch shape: targets per lane (1 to 4; 0 = automatic) and
 * source splits (0 = automatic).  Applies to later solves and steps. */
int lpm_bve_set_launch(lpm_bve_ctx* ctx, int targets_per_lane, int source_splits);

/* FP32C and FP64 Lagrangian-scheme sums: source tiles whose centre lies
 * within `factor` tile radii of a target use the explicit |y - x|^2 form,
 * farther tiles the expanded (1 - x.c) - x.d form (fewer operations; its
 * cancellation is bounded by ((factor+1)/(factor-1))^2).  Default 4; a tile
 * holding a target's own source always uses the explicit form. */
int lpm_bve_set_near_factor(lpm_bve_ctx* ctx, double factor);

/* Stage source tiles through LDS (1, default) or read them through the
 * scalar cache (0); both precisions. */
int lpm_bve_set_source_staging(lpm_bve_ctx* ctx, int lds);

/* The context's HIP stream (hipStream_t) and device id. */
int lpm_bve_stream(lpm_bve_ctx* ctx, void** stream, int* device);

/* Device time of the most recent pair-sum kernel launches queued by this
 * context (HIP events on the context's stream): launches, mean milliseconds. */
int lpm_bve_last_pair_timing(lpm_bve_ctx* ctx, int* launches, double* mean_ms);

/* Record HIP events around every pair-sum kernel this context queues (1) or
 * stop (0).  While on, steps run without graph replay. */
int lpm_bve_enable_timing(lpm_bve_ctx* ctx, int on);

/* Device memory a context needs for nv vertices and nactive active faces
 * at the default launch shape (state, source records, split partials).
 * LPM_ERR_ARG for sizes lpm_bve_set_state would reject. */
int lpm_bve_state_bytes(int64_t nv, int64_t nactive, int64_t* bytes);

/* ------------------------------------------------------------------ */
/* No counterpart in the reference (the batching form's autotuner).    */
/* Launch shapes keyed on the number of targets N (vertices + active   */
/* faces).  A context whose launch is automatic (lpm_bve_set_launch    */
/* with targets_per_lane 0, the default) uses the rule for its N, so a */
/* problem alone and inside an ensemble is decomposed identically.     */
/* ------------------------------------------------------------------ */
int lpm_shape_rule_set(int64_t ntargets, int targets_per_lane, int source_splits);
int lpm_shape_rule_get(int64_t ntargets, int* targets_per_lane, int* source_splits);
int lpm_shape_rules_clear(void);

/* Time one evaluation per candidate shape on the context's problem and
 * record the fastest as the rule for its N.  Outputs may be NULL. */
int lpm_bve_autotune(lpm_bve_ctx* ctx, int prec, int scheme, int reps, int* best_targets_per_lane,
                     int* best_source_splits, double* best_ms);

/* ------------------------------------------------------------------ */
/* No counterpart in the reference (the batching form).                */
/* Ensemble: independent problems (contexts, possibly of different     */
/* mesh levels) advanced together.  Each of the four evaluations of a  */
/* step runs as ONE persistent pair kernel over the flattened list of  */
/* (member, target block, source split) items, longest first (one per  */
/* targets-per-lane count), with one stage, one source-record and one  */
/* finalize launch over all members; each member keeps the             */
/* decomposition it has alone and writes its own partial slots, so its */
/* results are bit-identical alone and in the ensemble.  Members are   */
/* borrowed, must stay alive while they belong to the ensemble, and    */
/* must not be stepped on their own concurrently.                      */
/* ------------------------------------------------------------------ */
typedef struct lpm_bve_ensemble lpm_bve_ensemble;

int lpm_ens_create(int device, lpm_bve_ensemble** ens);
int lpm_ens_destroy(lpm_bve_ensemble* ens);
int lpm_ens_add(lpm_bve_ensemble* ens, lpm_bve_ctx* member);
int lpm_ens_remove(lpm_bve_ensemble* ens, lpm_bve_ctx* member);
int lpm_ens_size(lpm_bve_ensemble* ens, int* nmembers);
/* Workgroups per CU of the persistent kernel (0 = automatic: the
 * lpm_ens_autotune rule for the ensemble's total target count N, else the
 * occupancy limit). */
int lpm_ens_set_blocks_per_cu(lpm_bve_ensemble* ens, int blocks_per_cu);
/* No counterpart in the reference (the batching form's autotuner).
 * Time one evaluation of every member at its current state per candidate
 * persistent grid -- 1 up to the occupancy limit of workgroups (256 lanes,
 * 4 waves each) per CU -- and record the fewest workgroups per CU within
 * 0.5 % of the fastest as the process-wide rule for ensembles of this total
 * target count N (the members' targets summed).
 * Results do not depend on the grid.  ms_by_blocks[k] (k < nms; NULL when
 * nms is 0) receives the mean time at k + 1 workgroups per CU, -1 where not
 * tried.  The states are left alone; each member's next step recomputes its
 * entry velocity.  Outputs may be NULL. */
int lpm_ens_autotune(lpm_bve_ensemble* ens, int prec, int scheme, int reps, int* best_blocks_per_cu,
                     double* best_ms, double* ms_by_blocks, int nms);
/* Forget every ensemble grid rule. */
int lpm_ens_rules_clear(void);
/* No counterpart in the reference (the batching form's autotuner).
 * Launch shapes tuned for the ensemble: for the member sizes N that carry the
 * most pair work, largest first, until `work_fraction` of the ensemble's pair
 * work is covered, time one evaluation of every member at its state per
 * candidate shape of that N (targets per lane x source splits) and record
 * the fastest as the process-wide rule for N (lpm_shape_rule_get).  A
 * context of that N run alone takes the same rule, so results stay
 * bit-identical alone and in the ensemble.  Members must have automatic
 * launch shapes.  *ntuned (may be NULL) receives the number of sizes tuned;
 * each member's next step recomputes its entry velocity. */
int lpm_ens_autotune_shapes(lpm_bve_ensemble* ens, int prec, int scheme, int reps, double work_fraction,
                            int* ntuned);
/* `nsteps` RK4 steps of every member (same dt, omega, scheme, precision). */
int lpm_ens_advance(lpm_bve_ensemble* ens, double dt, double omega, int scheme, int prec, int nsteps,
                    int use_graph);
int lpm_ens_synchronize(lpm_bve_ensemble* ens);
/* Non-blocking: *done = 1 when the steps queued by the ensemble have finished. */
int lpm_ens_query(lpm_bve_ensemble* ens, int* done);
/* The ensemble's stream (hipStream_t), for members created on it. */
int lpm_ens_stream(lpm_bve_ensemble* ens, void** stream);

#ifdef __cplusplus
}
#endif

#endif /* LPM_BVE_H */
