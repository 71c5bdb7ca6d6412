"""Host-side mirror of the reference's BVE interface over the C ABI.

Names, argument meaning and call order follow the reference:

    seed = MeshSeed(IcosTriSphereSeed)                      # src/LpmMeshSeed.hpp:181-224
    nv, ne, nf = seed.setMaxAllocations(depth)              # src/LpmMeshSeed.cpp:296-305
    sphere = BVESphere(nv, ne, nf)                          # src/LpmBVESphere.hpp:40
    sphere.treeInit(depth, seed)                            # src/LpmPolyMesh2d.hpp:171
    sphere.set_omega(omega)                                 # src/LpmBVESphere.cpp:115-124
    sphere.init_vorticity(RossbyHaurwitzWave(4, w, K))      # src/LpmBVESphere.cpp:126-169
    solver = BVERK4(dt, Omega); solver.init(nv, nf)         # src/LpmBVERK4.hpp:29-34
    solver.advance_timestep(sphere.physVerts.crds, sphere.relVortVerts, sphere.velocityVerts,
                            sphere.physFaces.crds, sphere.relVortFaces, sphere.velocityFaces,
                            sphere.faces.area, sphere.faces.mask)

The particle state lives in device memory (one lpm_bve_ctx per sphere);
the view objects name fields of that state, and ``updateHost`` copies them
back, as the reference's host mirrors do.  All arithmetic runs in
liblpm_bve.so (host mesh build, host IC evaluation, device pair sums).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import PREC_FP32C, PREC_FP64, SCHEME_LAGR, SCHEME_REF, call

PI = 3.1415926535897932384626433832795027975


# ---------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------
class IcosTriSphereSeed:
    """src/LpmMeshSeed.hpp:132-147."""
    nverts, nfaces, nedges, nfaceverts = 12, 20, 30, 3
    kind = _lib.SEED_ICOS_TRI

    @staticmethod
    def idString() -> str:
        return "IcosTriSphereSeed"

    @staticmethod
    def faceStr() -> str:
        return "tri"

    @staticmethod
    def nVerticesAtTreeLevel(lev: int) -> int:
        return 2 + 10 * 4 ** lev

    @staticmethod
    def nFacesAtTreeLevel(lev: int) -> int:
        return 20 * 4 ** lev


class CubedSphereSeed:
    """src/LpmMeshSeed.hpp:111-126: quadrilateral faces; a divided face's
    centre becomes its children's shared corner."""
    nverts, nfaces, nedges, nfaceverts = 8, 6, 12, 4
    kind = _lib.SEED_CUBED_SPHERE

    @staticmethod
    def idString() -> str:
        return "CubedSphereSeed"

    @staticmethod
    def faceStr() -> str:
        return "quad"

    @staticmethod
    def nVerticesAtTreeLevel(lev: int) -> int:
        return 2 + 6 * 4 ** lev

    @staticmethod
    def nFacesAtTreeLevel(lev: int) -> int:
        return 6 * 4 ** lev


class MeshSeed:
    """MeshSeed<SeedType>: the built-in seed of ``seed_type`` (IcosTriSphereSeed
    or CubedSphereSeed), or a seed file in the reference's format when
    ``path`` is given."""

    def __init__(self, seed_type=IcosTriSphereSeed, path: str | None = None):
        if seed_type not in (IcosTriSphereSeed, CubedSphereSeed):
            raise NotImplementedError("the spherical seeds are IcosTriSphereSeed and CubedSphereSeed")
        self.seed_type = seed_type
        self.path = path

    @property
    def kind(self) -> int:
        return self.seed_type.kind

    def idString(self) -> str:
        return self.seed_type.idString()

    def setMaxAllocations(self, lev: int):
        nv, ne, nf = C.c_int64(), C.c_int64(), C.c_int64()
        call("lpm_seed_alloc", self.kind, lev, C.byref(nv), C.byref(ne), C.byref(nf))
        return nv.value, ne.value, nf.value


class TreeMesh:
    """Arrays of PolyMesh2d after treeInit (host copies)."""

    def __init__(self, level: int, seed: MeshSeed | None = None):
        seed = seed or MeshSeed()
        nv, ne, nf = seed.setMaxAllocations(level)
        self.level = level
        self.vx = np.zeros((nv, 3))
        self.vlag = np.zeros((nv, 3))
        self.fx = np.zeros((nf, 3))
        self.flag = np.zeros((nf, 3))
        self.farea = np.zeros(nf)
        self.fmask = np.zeros(nf, np.uint8)
        self.seed_type = seed.seed_type
        self.seed_kind = seed.kind
        self.seed_path = seed.path.encode() if seed.path else None
        self.fverts = np.zeros((nf, seed.seed_type.nfaceverts), np.int32)
        self.fparent = np.zeros(nf, np.int32)
        counts = np.zeros(4, np.int64)
        call("lpm_seed_tree_init", seed.kind, level, seed.path.encode() if seed.path else None,
             *(_lib.ptr(a) for a in (self.vx, self.vlag, self.fx, self.flag, self.farea, self.fmask, self.fverts,
                                     self.fparent, counts)))
        self.nverts, self.nedges, self.nfaces, self.nleaves = (int(c) for c in counts)

    def leaf_index(self):
        return np.nonzero(self.fmask == 0)[0]

    @classmethod
    def from_nc(cls, path: str) -> "TreeMesh":
        """PolyMesh2d(const PolyMeshReader&) (src/LpmPolyMesh2d.cpp:20-28): the
        arrays of a mesh file (write_mesh_nc or a checkpoint with topology)."""
        counts = np.zeros(4, np.int64)
        call("lpm_mesh_read_nc", path.encode(), *([None] * 8), _lib.ptr(counts))
        nv, ne, nf, nl = (int(c) for c in counts)
        seed_type = CubedSphereSeed if nv == nl + 2 else IcosTriSphereSeed
        m = cls.__new__(cls)
        m.seed_type = seed_type
        m.seed_kind = _lib.SEED_CUBED_SPHERE if seed_type is CubedSphereSeed else _lib.SEED_ICOS_TRI
        m.seed_path = None
        m.vx, m.vlag = np.zeros((nv, 3)), np.zeros((nv, 3))
        m.fx, m.flag = np.zeros((nf, 3)), np.zeros((nf, 3))
        m.farea, m.fmask = np.zeros(nf), np.zeros(nf, np.uint8)
        m.fverts, m.fparent = np.zeros((nf, seed_type.nfaceverts), np.int32), np.zeros(nf, np.int32)
        call("lpm_mesh_read_nc", path.encode(),
             *(_lib.ptr(a) for a in (m.vx, m.vlag, m.fx, m.flag, m.farea, m.fmask, m.fverts, m.fparent, counts)))
        m.nverts, m.nedges, m.nfaces, m.nleaves = (int(c) for c in counts)
        m.level = next((lev for lev in range(13) if seed_type.nVerticesAtTreeLevel(lev) == nv), -1)
        return m


def write_mesh_vtk(level: int, path: str, fmt: int = _lib.VTK_LEGACY_ASCII, seed: MeshSeed | None = None):
    """PolyMesh2d::outputVtk (src/LpmPolyMesh2d.cpp:79-85) of the level-`level`
    treeInit mesh: vertices, leaf faces and their areas."""
    seed = seed or MeshSeed()
    call("lpm_seed_mesh_write_vtk", seed.kind, int(level), seed.path.encode() if seed.path else None, path.encode(),
         int(fmt))


def write_mesh_nc(level: int, path: str, seed: MeshSeed | None = None):
    """NcWriter::writePolymesh (src/LpmNetCDF_Impl.hpp:46-218) of the level-`level`
    treeInit mesh, as a netCDF classic (64-bit offset) file."""
    seed = seed or MeshSeed()
    call("lpm_seed_mesh_write_nc", seed.kind, int(level), seed.path.encode() if seed.path else None, path.encode())


# ---------------------------------------------------------------------------
# initial conditions: VorticityInitialCondition::eval (src/LpmVorticityGallery.hpp:13-40)
# ---------------------------------------------------------------------------
class VorticityInitialCondition:
    """Composite field: solid-body rotation, Rossby-Haurwitz wave, Gaussian vortices
    (evaluated by lpm_ic_eval in liblpm_bve.so)."""

    def __init__(self, sbr: float = 0.0, rh_n: int = 0, rh_w: float = 0.0, rh_K: float = 0.0,
                 centers=None, amps=None, sigma: float = 1.0):
        self.sbr, self.rh_n, self.rh_w, self.rh_K, self.sigma = sbr, rh_n, rh_w, rh_K, sigma
        self.centers = None if centers is None else np.ascontiguousarray(centers, np.float64).reshape(-1, 3)
        self.amps = None if amps is None else np.ascontiguousarray(amps, np.float64)

    def with_vortices(self, seed: int, nvort: int, amp_scale: float, sigma: float):
        c, a = gauss_params(seed, nvort, amp_scale)
        self.centers, self.amps, self.sigma = c, a, sigma
        return self

    def eval_many(self, xyz) -> np.ndarray:
        xyz = np.ascontiguousarray(xyz, np.float64)
        out = np.zeros(xyz.shape[0])
        nv = 0 if self.amps is None else len(self.amps)
        call("lpm_ic_eval", xyz, xyz.shape[0], out, float(self.sbr), int(self.rh_n), float(self.rh_w),
             float(self.rh_K), nv, _lib.ptr(self.centers) if nv else None, _lib.ptr(self.amps) if nv else None,
             float(self.sigma))
        return out

    def eval(self, x: float, y: float, z: float) -> float:
        return float(self.eval_many(np.array([[x, y, z]]))[0])


class SolidBodyRotation(VorticityInitialCondition):
    """zeta = 2 * OMEGA * z, OMEGA = 2 pi (src/LpmVorticityGallery.hpp:25-40)."""
    OMEGA = 2 * PI

    def __init__(self):
        super().__init__(sbr=1.0)


class RossbyHaurwitzWave(VorticityInitialCondition):
    """zeta = 2 w sin(lat) - K (n+1)(n+2) cos^n(lat) sin(lat) cos(n lon)."""

    def __init__(self, n: int, w: float, K: float):
        super().__init__(rh_n=n, rh_w=w, rh_K=K)

    def phase_speed(self, omega: float) -> float:
        n, w = self.rh_n, self.rh_w
        return (n * (n + 3) * w - 2 * omega) / ((n + 1) * (n + 2))


def gauss_params(seed: int, nvort: int, amp_scale: float = 1.0):
    c = np.zeros((nvort, 3))
    a = np.zeros(nvort)
    if nvort:
        call("lpm_ic_gauss_params", seed, nvort, amp_scale, c, a)
    return c, a


def area_mean(fzeta, farea, fmask) -> float:
    m = C.c_double()
    call("lpm_area_mean", np.ascontiguousarray(fzeta, np.float64), np.ascontiguousarray(farea, np.float64),
         np.ascontiguousarray(fmask, np.uint8), len(fzeta), C.byref(m))
    return m.value


# ---------------------------------------------------------------------------
# device state
# ---------------------------------------------------------------------------
class ErrNorms:
    """ErrNorms (src/LpmErrorNorms.hpp:81-130, src/LpmErrorNorms.cpp:7-70):
    relative l1, l2 and linf errors with weights (areas), computed by the
    library on the host in index order.  vector(): fields [n][3], the error
    appx - exact is formed (and kept in .err); scalar(): a precomputed
    error field, with the reference's signed maxima in linf."""

    def __init__(self, l1: float, l2: float, linf: float, err=None):
        self.l1, self.l2, self.linf = float(l1), float(l2), float(linf)
        self.err = err

    @classmethod
    def vector(cls, appx, exact, wt) -> "ErrNorms":
        a = np.ascontiguousarray(appx, np.float64).reshape(-1, 3)
        x = np.ascontiguousarray(exact, np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(wt, np.float64)
        if a.shape != x.shape or w.shape != (a.shape[0],):
            raise ValueError("ErrNorms.vector: appx and exact [n][3], wt [n]")
        err = np.empty_like(a)
        out = np.zeros(3)
        _lib.call("lpm_err_norms_vec", err.ctypes.data, a, x, w, a.shape[0], out)
        return cls(*out, err=err)

    @classmethod
    def scalar(cls, err, exact, wt) -> "ErrNorms":
        e = np.ascontiguousarray(err, np.float64)
        x = np.ascontiguousarray(exact, np.float64)
        w = np.ascontiguousarray(wt, np.float64)
        if e.shape != x.shape or e.shape != w.shape or e.ndim != 1:
            raise ValueError("ErrNorms.scalar: err, exact and wt [n]")
        out = np.zeros(3)
        _lib.call("lpm_err_norms_scalar", e, x, w, e.shape[0], out)
        return cls(*out)

    def infoString(self, label: str = "", tab_level: int = 0) -> str:
        return "\t" * tab_level + f"{label} ErrNorms: l1 = {self.l1:g} l2 = {self.l2:g} linf = {self.linf:g}\n"


def velocity_tangent(x, u) -> np.ndarray:
    """SphereVelocityTangentTestFunctor (src/LpmSphereTestKernels.hpp:11-25): x_i . u_i."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    if x.shape != u.shape:
        raise ValueError("velocity_tangent: x and u [n][3]")
    out = np.zeros(x.shape[0])
    _lib.call("lpm_velocity_tangent", x, u, x.shape[0], out)
    return out


class Context:
    """One lpm_bve_ctx: a particle state resident in device memory."""

    def __init__(self, device: int = 0, stream=None):
        """stream: a hipStream_t (int or c_void_p) to queue work on instead of
        the context's own stream; it must outlive the context."""
        h = C.c_void_p()
        if stream is None:
            call("lpm_bve_create", device, C.byref(h))
        else:
            call("lpm_bve_create_on_stream", device, C.c_void_p(int(stream.value if hasattr(stream, "value")
                                                                    else stream)), C.byref(h))
        self.h = h
        self.device = device
        self.nv = self.nf = 0

    def close(self):
        if self.h:
            _lib.lib().lpm_bve_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, vx, vzeta, fx, fzeta, farea, fmask):
        vx = np.ascontiguousarray(vx, np.float64).reshape(-1, 3)
        fx = np.ascontiguousarray(fx, np.float64).reshape(-1, 3)
        self.nv, self.nf = vx.shape[0], fx.shape[0]
        call("lpm_bve_set_state", self.h, self.nv, vx, np.ascontiguousarray(vzeta, np.float64), self.nf, fx,
             np.ascontiguousarray(fzeta, np.float64), np.ascontiguousarray(farea, np.float64),
             np.ascontiguousarray(fmask, np.uint8))

    def set_state_level(self, level: int, ic: "VorticityInitialCondition", subtract_mean: bool = False,
                        seed_kind: int = _lib.SEED_ICOS_TRI):
        """State of the level-`level` tree of seed `seed_kind` (every tree face;
        icosahedral by default, SEED_CUBED_SPHERE for the cubed sphere) from the
        device-resident mesh cache, with `ic` evaluated on the device."""
        nvort = 0 if ic.amps is None else len(ic.amps)
        centers = np.ascontiguousarray(ic.centers, np.float64) if nvort else None
        amps = np.ascontiguousarray(ic.amps, np.float64) if nvort else None
        p = _lib.IcParams(float(ic.sbr), int(ic.rh_n), float(ic.rh_w), float(ic.rh_K), nvort,
                          centers.ctypes.data if nvort else None, amps.ctypes.data if nvort else None,
                          float(ic.sigma), int(subtract_mean))
        call("lpm_bve_set_state_seed_level", self.h, int(seed_kind), int(level), C.byref(p))
        sizes = self.sizes()
        self.nv, self.nf = sizes[0], sizes[1]

    def set_velocity(self, vu, fu):
        call("lpm_bve_set_velocity", self.h, np.ascontiguousarray(vu, np.float64),
             np.ascontiguousarray(fu, np.float64))

    def solve(self, prec: int = PREC_FP64, scheme: int = SCHEME_REF, want_psi: bool = False):
        call("lpm_bve_solve", self.h, prec, scheme, int(want_psi))

    def streamfn(self, prec: int = PREC_FP64, scheme: int = SCHEME_REF):
        """Stream function alone at the current state (the reference driver's
        post-timestep solve); the held velocity is left as it is."""
        call("lpm_bve_streamfn", self.h, prec, scheme)

    def advance(self, dt: float, omega: float, scheme: int = SCHEME_REF, prec: int = PREC_FP64, nsteps: int = 1,
                use_graph: bool = True):
        call("lpm_bve_advance", self.h, dt, omega, scheme, prec, nsteps, int(use_graph))

    def set_launch(self, targets_per_lane: int, source_splits: int = 0):
        call("lpm_bve_set_launch", self.h, targets_per_lane, source_splits)

    def set_source_staging(self, lds: bool):
        call("lpm_bve_set_source_staging", self.h, int(lds))

    def set_near_factor(self, factor: float):
        call("lpm_bve_set_near_factor", self.h, factor)

    def enable_timing(self, on: bool):
        call("lpm_bve_enable_timing", self.h, int(on))

    def pair_timing(self):
        n, ms = C.c_int(), C.c_double()
        call("lpm_bve_last_pair_timing", self.h, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def synchronize(self):
        call("lpm_bve_synchronize", self.h)

    def done(self) -> bool:
        """True when all work queued on the context has finished (non-blocking)."""
        d = C.c_int()
        call("lpm_bve_query", self.h, C.byref(d))
        return bool(d.value)

    def sizes(self):
        s = np.zeros(4, np.int64)
        call("lpm_bve_sizes", self.h, s)
        return tuple(int(v) for v in s)

    def circulation(self) -> float:
        c = C.c_double()
        call("lpm_bve_circulation", self.h, C.byref(c))
        return c.value

    def checkpoint(self, path: str, time: float = 0.0, omega: float = 0.0, level: int = -1):
        """Write the state to a netCDF file that restore() resumes from bit for
        bit (mesh variables, BVESphere fields, time and rotation rate).  level:
        the tree the state was built on, for the full mesh topology (-1: the
        state's own level when built by level, else particles only)."""
        call("lpm_bve_checkpoint", self.h, path.encode(), int(level), float(time), float(omega))

    def restore(self, path: str):
        """Load a checkpoint; returns its (time, omega)."""
        t, om = C.c_double(), C.c_double()
        call("lpm_bve_restore", self.h, path.encode(), C.byref(t), C.byref(om))
        sizes = self.sizes()
        self.nv, self.nf = sizes[0], sizes[1]
        return t.value, om.value

    def write_vtk(self, path: str, omega: float = 0.0, level: int = -1, fmt: int = _lib.VTK_LEGACY_ASCII,
                  arrays=None):
        """VTK polydata of the state on its tree's leaf faces (fmt: VTK_LEGACY_ASCII,
        VTK_LEGACY_BINARY, VTK_XML_ASCII, VTK_XML_BINARY).  arrays: more fields as
        (name, on_faces, values) with values [nverts] / [nverts][k] at vertices or
        [nfaces] / [nfaces][k] at faces (tree order)."""
        arrays = arrays or []
        keep = []
        items = (_lib.VtkArray * max(1, len(arrays)))()
        for i, (name, on_faces, values) in enumerate(arrays):
            a = np.ascontiguousarray(values, np.float64)
            n = self.nf if on_faces else self.nv
            if a.shape[0] != n:
                raise ValueError(f"array {name}: {a.shape[0]} rows, expected {n}")
            keep.append(a)
            items[i] = _lib.VtkArray(name.encode(), 1 if a.ndim == 1 else int(a.shape[1]), int(bool(on_faces)),
                                     a.ctypes.data)
        call("lpm_bve_write_vtk_arrays", self.h, path.encode(), int(level), float(omega), int(fmt), items,
             len(arrays))

    def get_state(self):
        nv, nf = self.nv, self.nf
        out = dict(vx=np.zeros((nv, 3)), vzeta=np.zeros(nv), vu=np.zeros((nv, 3)), vpsi=np.zeros(nv),
                   fx=np.zeros((nf, 3)), fzeta=np.zeros(nf), fu=np.zeros((nf, 3)), fpsi=np.zeros(nf))
        call("lpm_bve_get_state", self.h, *(_lib.ptr(out[k]) for k in
                                            ("vx", "vzeta", "vu", "vpsi", "fx", "fzeta", "fu", "fpsi")))
        return out


class DeviceView:
    """Handle naming one field of a sphere's resident state (the role of a
    Kokkos View handle in BVERK4::advance_timestep's signature)."""

    def __init__(self, sphere: "BVESphere", name: str):
        self.sphere, self.name = sphere, name


class _Coords:
    def __init__(self, sphere, which):
        self.crds = DeviceView(sphere, which + "x")


class _Faces:
    def __init__(self, sphere):
        self.area = DeviceView(sphere, "farea")
        self.mask = DeviceView(sphere, "fmask")


class BVESphere:
    """BVESphere<Seed> (src/LpmBVESphere.hpp:18-80; the icosahedral and the
    cubed-sphere seeds, src/LpmBVESphere.cpp:201-202) with fields resident on
    device `device`; the seed comes with treeInit."""

    def __init__(self, nmaxverts: int, nmaxedges: int, nmaxfaces: int, nq: int = 0, device: int = 0):
        self.nmax = (nmaxverts, nmaxedges, nmaxfaces)
        self.Omega = 2 * PI  # src/LpmBVESphere.cpp:24
        self._omg_set = False
        self.t = 0.0
        self.ctx = Context(device)
        self.mesh: TreeMesh | None = None
        self.physVerts = _Coords(self, "v")
        self.physFaces = _Coords(self, "f")
        self.faces = _Faces(self)
        self.relVortVerts = DeviceView(self, "vzeta")
        self.relVortFaces = DeviceView(self, "fzeta")
        self.velocityVerts = DeviceView(self, "vu")
        self.velocityFaces = DeviceView(self, "fu")
        self.host = {}
        # passive tracers (src/LpmBVESphere.cpp:36-58): host arrays carried by the
        # particles, written with the state by outputVtk
        self.tracer_names: list[str] = []
        self.tracer_verts: list[np.ndarray] = []
        self.tracer_faces: list[np.ndarray] = []
        for k in range(nq):
            self.create_tracer(f"tracer{k}")

    def create_tracer(self, name: str) -> int:
        """BVESphere::create_tracer: a new zero tracer named `name`; returns its index."""
        self.tracer_names.append(name)
        self.tracer_verts.append(np.zeros(self.nmax[0]))
        self.tracer_faces.append(np.zeros(self.nmax[2]))
        return len(self.tracer_names) - 1

    # PolyMesh2d::treeInit
    def treeInit(self, depth: int, seed: MeshSeed):
        self.mesh = TreeMesh(depth, seed)
        if self.mesh.nverts > self.nmax[0] or self.mesh.nfaces > self.nmax[2]:
            raise ValueError("treeInit: mesh exceeds the allocation sizes")

    def nvertsHost(self) -> int:
        return self.mesh.nverts

    def nfacesHost(self) -> int:
        return self.mesh.nfaces

    def set_omega(self, omg: float):
        if self._omg_set:
            print(f"BVESphere::set_omega warning: omega = {self.Omega} already set.")
        else:
            self.Omega = omg
            self._omg_set = True

    def init_vorticity(self, ic: VorticityInitialCondition, subtract_mean: bool = False, prec: int = PREC_FP64,
                       scheme: int = SCHEME_REF):
        """Evaluate the IC at every particle (host), set absVort = zeta + 2 Omega z,
        upload, and compute stream function and velocity on the device."""
        m = self.mesh
        vz = ic.eval_many(m.vx)
        fz = ic.eval_many(m.fx)
        if subtract_mean:
            mu = area_mean(fz, m.farea, m.fmask)
            vz = vz - mu
            fz = fz - mu
        self.host["absVortVerts"] = vz + 2 * self.Omega * m.vx[:, 2]
        self.host["absVortFaces"] = fz + 2 * self.Omega * m.fx[:, 2]
        self.ctx.set_state(m.vx, vz, m.fx, fz, m.farea, m.fmask)
        self.ctx.solve(prec, scheme, want_psi=True)

    def solve_streamfn(self, prec: int = PREC_FP64, scheme: int = SCHEME_REF):
        """Stream function at the current state (tests/LpmBVETest.cpp:155-159);
        the velocity, the next step's entry velocity, is left as it is."""
        self.ctx.streamfn(prec, scheme)

    def updateHost(self):
        s = self.ctx.get_state()
        self.host.update(s)
        return s

    def outputVtk(self, fname: str, arrays=None):
        """BVESphere::outputVtk (src/LpmBVESphere.cpp:90-113): legacy VTK polydata
        of the leaf faces; a .vtp name writes the XML layout of
        Polymesh2dVtkInterface::write with addFieldsToVtk instead.  The tracers
        follow the state's fields; arrays: more fields after them, as
        Context.write_vtk takes them."""
        fmt = _lib.VTK_XML_BINARY if fname.endswith(".vtp") else _lib.VTK_LEGACY_ASCII
        nv, nf = self.mesh.nverts, self.mesh.nfaces
        tracers = []
        for name, tv, tf in zip(self.tracer_names, self.tracer_verts, self.tracer_faces):
            tracers += [(name, False, tv[:nv]), (name, True, tf[:nf])]
        self.ctx.write_vtk(fname, self.Omega, self.mesh.level, fmt, tracers + list(arrays or []))

    def checkpoint(self, fname: str):
        """The mesh, fields, time and rotation rate in one netCDF file."""
        self.ctx.checkpoint(fname, self.t, self.Omega, self.mesh.level)

    def restore(self, fname: str):
        """Resume from checkpoint(); the tree is rebuilt from its level."""
        self.t, self.Omega = self.ctx.restore(fname)
        self._omg_set = True
        nv = self.ctx.nv
        for seed_type in (IcosTriSphereSeed, CubedSphereSeed):
            level = next((lev for lev in range(13) if seed_type.nVerticesAtTreeLevel(lev) == nv), None)
            if level is not None:
                break
        if self.mesh is None or self.mesh.level != level or self.mesh.seed_type is not seed_type:
            self.mesh = TreeMesh(level, MeshSeed(seed_type))

    def avg_mesh_size_radians(self) -> float:
        return math.sqrt(4 * PI / self.nfacesHost())

    def avg_mesh_size_degrees(self) -> float:
        return math.degrees(self.avg_mesh_size_radians())

    def circulation(self) -> float:
        return self.ctx.circulation()


class BVERK4:
    """BVERK4 (src/LpmBVERK4.hpp:13-65): classical RK4 over the sphere's
    resident state.  scheme/prec select the reference step (REF, FP64) or the
    renormalised Lagrangian step in FP64 or FP32C."""

    def __init__(self, timestep: float, omg: float, scheme: int = SCHEME_REF, prec: int = PREC_FP64,
                 use_graph: bool = True):
        self.dt, self.Omega = timestep, omg
        self.scheme, self.prec, self.use_graph = scheme, prec, use_graph
        self.nverts = self.nfaces = 0

    def init(self, nv: int, nf: int):
        self.nverts, self.nfaces = nv, nf

    def advance_timestep(self, vx, vzeta, vvel, fx, fzeta, fvel, fa, fm, nsteps: int = 1):
        views = (vx, vzeta, vvel, fx, fzeta, fvel, fa, fm)
        names = ("vx", "vzeta", "vu", "fx", "fzeta", "fu", "farea", "fmask")
        sphere = vx.sphere
        for v, n in zip(views, names):
            if not isinstance(v, DeviceView) or v.sphere is not sphere or v.name != n:
                raise ValueError("advance_timestep: arguments must be the sphere's "
                                 "(vx, vzeta, vvel, fx, fzeta, fvel, area, mask) views, in that order")
        if sphere.mesh.nverts != self.nverts or sphere.mesh.nfaces != self.nfaces:
            raise ValueError("advance_timestep: init(nverts, nfaces) does not match the sphere")
        sphere.ctx.advance(self.dt, self.Omega, self.scheme, self.prec, nsteps, self.use_graph)
        sphere.t += nsteps * self.dt


# ---------------------------------------------------------------------------
# SpherePoisson (src/LpmSpherePoisson.hpp:359-520)
# ---------------------------------------------------------------------------
def poisson_y54(xyz):
    """f = Y_5^4 and the velocity RH54Velocity at points xyz [n][3]."""
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    f = np.zeros(xyz.shape[0])
    u = np.zeros_like(xyz)
    call("lpm_poisson_y54", xyz, xyz.shape[0], _lib.ptr(f), _lib.ptr(u))
    return f, u


class SpherePoisson:
    """-lap(psi) = f with f = Y_5^4 on the sphere (exact psi = f/30, exact
    velocity RH54Velocity), solved by the direct sums on the GPU (FP64, literal
    1 - x.y); the reference's L-infinity errors over vertices and active faces."""

    def __init__(self, nmaxverts: int, nmaxedges: int, nmaxfaces: int, device: int = 0):
        self.nmax = (nmaxverts, nmaxedges, nmaxfaces)
        self.device = device
        self.linf_verts = self.linf_faces = self.linf_uverts = self.linf_ufaces = 0.0

    def treeInit(self, depth: int, seed: MeshSeed | None = None):
        self.mesh = TreeMesh(depth, seed)
        if self.mesh.nverts > self.nmax[0] or self.mesh.nfaces > self.nmax[2]:
            raise ValueError("treeInit: not enough memory")

    def meshSize(self) -> float:
        return math.sqrt(4 * PI / self.mesh.nleaves)

    def init(self):
        m = self.mesh
        self.fverts, self.uvertsexact = poisson_y54(m.vx)
        self.ffaces, self.ufacesexact = poisson_y54(m.fx)
        self.psiexactverts = self.fverts / 30.0
        self.psiexactfaces = self.ffaces / 30.0

    def solve(self, nthreads: int = 0, quiet: bool = False):
        m = self.mesh
        c = Context(self.device)
        try:
            c.set_state(m.vx, self.fverts, m.fx, self.ffaces, m.farea, m.fmask)
            c.solve(PREC_FP64, SCHEME_REF, want_psi=True)
            s = c.get_state()
        finally:
            c.close()
        self.psiverts, self.psifaces, self.uverts, self.ufaces = s["vpsi"], s["fpsi"], s["vu"], s["fu"]
        lv = m.fmask == 0
        self.everts = np.abs(self.psiverts - self.psiexactverts)
        self.efaces = np.where(lv, np.abs(self.psifaces - self.psiexactfaces), 0.0)
        self.uerrverts = np.abs(self.uverts - self.uvertsexact)
        self.uerrfaces = np.where(lv[:, None], np.abs(self.ufaces - self.ufacesexact), 0.0)
        self.linf_verts = float(self.everts.max())
        self.linf_faces = float(self.efaces.max())
        self.linf_uverts = float(np.sum(self.uerrverts ** 2, axis=1).max())
        self.linf_ufaces = float(np.sum(self.uerrfaces ** 2, axis=1).max())
        if not quiet:
            print(f"{self.meshSize() * 180.0 / PI:g}:\n"
                  f"\tpotential: linf_verts = {self.linf_verts:g}, linf_faces = {self.linf_faces:g}\n"
                  f"\tvelocity:  linf_verts = {math.sqrt(self.linf_uverts):g}, "
                  f"linf_faces = {math.sqrt(self.linf_ufaces):g}")

    def outputVtk(self, fname: str):
        """SpherePoisson::outputVtk (src/LpmSpherePoisson.hpp:510-531): legacy VTK
        polydata of the leaf faces (XML for a name ending in .vtp) with f, psi,
        psi_exact, psi_error, velocity, velocity_exact and velocity_error at
        vertices and faces, cell "area" first."""
        m = self.mesh
        fields = [("f", self.fverts, self.ffaces), ("psi", self.psiverts, self.psifaces),
                  ("psi_exact", self.psiexactverts, self.psiexactfaces), ("psi_error", self.everts, self.efaces),
                  ("velocity", self.uverts, self.ufaces), ("velocity_exact", self.uvertsexact, self.ufacesexact),
                  ("velocity_error", self.uerrverts, self.uerrfaces)]
        rows = [(name, False, v) for name, v, _ in fields] + [(name, True, f) for name, _, f in fields]
        keep = []
        items = (_lib.VtkArray * len(rows))()
        for i, (name, on_faces, values) in enumerate(rows):
            a = np.ascontiguousarray(values, np.float64)
            keep.append(a)
            items[i] = _lib.VtkArray(name.encode(), 1 if a.ndim == 1 else int(a.shape[1]), int(on_faces), a.ctypes.data)
        fmt = _lib.VTK_XML_BINARY if fname.endswith(".vtp") else _lib.VTK_LEGACY_ASCII
        call("lpm_seed_mesh_write_vtk_arrays", m.seed_kind, m.level, m.seed_path, fname.encode(), fmt, items, len(rows))
