"""ctypes front end for the CPU oracle (liblpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py for expected values.  The product package never imports
this module.  Each wrapper names the reference function it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpm_oracle.so")
SEED_FIXTURE = os.path.join(os.path.dirname(HERE), "tests", "golden", "icosTriSphereSeed.dat")
CUBED_SEED_FIXTURE = os.path.join(os.path.dirname(HERE), "tests", "golden", "cubedSphereSeed.dat")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_seed_read.argtypes = [C.c_char_p, _dp, _ip, _ip, _ip]
        L.orc_icos_alloc.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_icos_tree_init.argtypes = [C.c_int, _dp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _bp, _ip, _ip, _ip]
        L.orc_cubed_seed_read.argtypes = [C.c_char_p, _dp, _ip, _ip, _ip]
        L.orc_cubed_alloc.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_cubed_tree_init.argtypes = [C.c_int, _dp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _bp, _ip, _ip, _ip]
        L.orc_velocity.argtypes = [_dp, _dp, C.c_int, _dp, _dp, _dp, _bp, C.c_int, C.c_int, C.c_int]
        L.orc_streamfn.argtypes = [_dp, _dp, C.c_int, _dp, _dp, _dp, _bp, C.c_int, C.c_int]
        L.orc_velocity_chord.argtypes = [_dp, _dp, C.c_int, _dp, _dp, _dp, _bp, C.c_int, C.c_int, C.c_int]
        L.orc_streamfn_chord.argtypes = [_dp, _dp, C.c_int, _dp, _dp, _dp, _bp, C.c_int, C.c_int]
        L.orc_velocity_ld.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _bp, C.c_int, C.c_int]
        L.orc_rk4_step.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                   _dp, _bp, C.c_void_p, C.c_void_p]
        L.orc_blas_update.argtypes = [_dp, _dp, C.c_double, _dp, C.c_int]
        L.orc_blas_scal.argtypes = [_dp, C.c_double, _dp, C.c_int]
        L.orc_rk4_combine.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_xoshiro_raw.argtypes = [_up, _up, C.c_int]
        L.orc_rng_draws.argtypes = [C.c_uint64, _dp, C.c_int, _dp, C.c_int]
        L.orc_gauss_params.argtypes = [C.c_uint64, C.c_int, C.c_double, _dp, _dp]
        L.orc_ic_eval.argtypes = [_dp, C.c_int, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                  C.c_double]
        L.orc_err_norms_vec.argtypes = [_dp, _dp, _dp, C.c_int, _dp]
        L.orc_err_norms_scalar.argtypes = [_dp, _dp, _dp, C.c_int, _dp]
        L.orc_sphharm54.argtypes = [_dp, C.c_int, _dp, _dp]
        for name in ("orc_tri_area", "orc_distance", "orc_latitude", "orc_longitude"):
            getattr(L, name).restype = C.c_double
        L.orc_tri_area.argtypes = [_dp, _dp, _dp]
        L.orc_polygon_area.argtypes = [_dp, _dp, C.c_int]
        L.orc_polygon_area.restype = C.c_double
        L.orc_midpoint.argtypes = [_dp, _dp, _dp]
        L.orc_barycenter.argtypes = [_dp, _dp, C.c_int]
        L.orc_cross.argtypes = [_dp, _dp, _dp]
        L.orc_distance.argtypes = [_dp, _dp]
        L.orc_latitude.argtypes = [_dp]
        L.orc_longitude.argtypes = [_dp]
        _lib = L
    return _lib


def read_seed(path: str = SEED_FIXTURE):
    """MeshSeed<IcosTriSphereSeed>::readfile (src/LpmMeshSeed.cpp:20-232)."""
    crds = np.zeros((32, 3))
    edges = np.zeros((30, 4), np.int32)
    fv = np.zeros((20, 3), np.int32)
    fe = np.zeros((20, 3), np.int32)
    rc = lib().orc_seed_read(path.encode(), crds, edges, fv, fe)
    if rc != 0:
        raise RuntimeError(f"seed read failed ({rc}) for {path}")
    return crds, edges, fv, fe


def icos_alloc(level: int):
    """setMaxAllocations (src/LpmMeshSeed.cpp:296-305): (nverts, nedges, nfaces)."""
    nv, ne, nf = C.c_int(), C.c_int(), C.c_int()
    lib().orc_icos_alloc(level, C.byref(nv), C.byref(ne), C.byref(nf))
    return nv.value, ne.value, nf.value


def read_cubed_seed(path: str = CUBED_SEED_FIXTURE):
    """MeshSeed<CubedSphereSeed>::readfile (src/LpmMeshSeed.cpp:20-232)."""
    crds = np.zeros((14, 3))
    edges = np.zeros((12, 4), np.int32)
    fv = np.zeros((6, 4), np.int32)
    fe = np.zeros((6, 4), np.int32)
    rc = lib().orc_cubed_seed_read(path.encode(), crds, edges, fv, fe)
    if rc != 0:
        raise RuntimeError(f"seed read failed ({rc}) for {path}")
    return crds, edges, fv, fe


def cubed_alloc(level: int):
    """MeshSeed<CubedSphereSeed>::setMaxAllocations (src/LpmMeshSeed.cpp:296-305, :369-379)."""
    nv, ne, nf = C.c_int(), C.c_int(), C.c_int()
    lib().orc_cubed_alloc(level, C.byref(nv), C.byref(ne), C.byref(nf))
    return nv.value, ne.value, nf.value


class Mesh:
    """PolyMesh2d<Seed>::treeInit (src/LpmPolyMesh2d.cpp:30-44) output, for the
    icosahedral triangle seed (kind "icos") or the cubed-sphere quadrilateral
    seed (kind "cubed")."""

    def __init__(self, level: int, seed=None, kind: str = "icos"):
        cubed = kind == "cubed"
        crds, edges, fv, fe = seed if seed is not None else (read_cubed_seed() if cubed else read_seed())
        nv, ne, nf = cubed_alloc(level) if cubed else icos_alloc(level)
        self.level = level
        self.kind = kind
        self.nfaceverts = 4 if cubed else 3
        self.vx = np.zeros((nv, 3))
        self.vlag = np.zeros((nv, 3))
        self.fx = np.zeros((nf, 3))
        self.flag = np.zeros((nf, 3))
        self.farea = np.zeros(nf)
        self.fmask = np.zeros(nf, np.uint8)
        self.fverts = np.zeros((nf, self.nfaceverts), np.int32)
        self.fparent = np.zeros(nf, np.int32)
        counts = np.zeros(4, np.int32)
        init = lib().orc_cubed_tree_init if cubed else lib().orc_icos_tree_init
        init(level, crds, edges, fv, fe, self.vx, self.vlag, self.fx, self.flag, self.farea, self.fmask,
             self.fverts, self.fparent, counts)
        self.nv, self.ne, self.nf, self.nleaves = (int(c) for c in counts)
        assert (self.nv, self.nf) == (nv, nf)

    @property
    def leaves(self):
        return np.nonzero(self.fmask == 0)[0]


def velocity(tx, fx, fzeta, farea, fmask, collocated: bool, stride: int = 1):
    """BVEVertexVelocity / BVEFaceVelocity (src/LpmBVEKernels.hpp:163-186, 285-305)."""
    tx = np.ascontiguousarray(tx, np.float64)
    u = np.zeros_like(tx)
    lib().orc_velocity(u, tx, tx.shape[0], np.ascontiguousarray(fx), np.ascontiguousarray(fzeta),
                       np.ascontiguousarray(farea), np.ascontiguousarray(fmask, np.uint8), fx.shape[0],
                       int(collocated), stride)
    return u


def streamfn(tx, fx, fzeta, farea, fmask, collocated: bool):
    """BVEVertexStreamFn / BVEFaceStreamFn (src/LpmBVEKernels.hpp:140-161, 263-283)."""
    tx = np.ascontiguousarray(tx, np.float64)
    psi = np.zeros(tx.shape[0])
    lib().orc_streamfn(psi, tx, tx.shape[0], np.ascontiguousarray(fx), np.ascontiguousarray(fzeta),
                       np.ascontiguousarray(farea), np.ascontiguousarray(fmask, np.uint8), fx.shape[0],
                       int(collocated))
    return psi


def velocity_chord(tx, fx, fzeta, farea, fmask, collocated: bool, stride: int = 1):
    """The same sums with the chord denominator |y - x|^2 / 2 (every `stride`-th
    target; the others stay zero)."""
    tx = np.ascontiguousarray(tx, np.float64)
    u = np.zeros_like(tx)
    lib().orc_velocity_chord(u, tx, tx.shape[0], np.ascontiguousarray(fx), np.ascontiguousarray(fzeta),
                             np.ascontiguousarray(farea), np.ascontiguousarray(fmask, np.uint8), fx.shape[0],
                             int(collocated), int(stride))
    return u


def streamfn_chord(tx, fx, fzeta, farea, fmask, collocated: bool):
    tx = np.ascontiguousarray(tx, np.float64)
    psi = np.zeros(tx.shape[0])
    lib().orc_streamfn_chord(psi, tx, tx.shape[0], np.ascontiguousarray(fx), np.ascontiguousarray(fzeta),
                             np.ascontiguousarray(farea), np.ascontiguousarray(fmask, np.uint8), fx.shape[0],
                             int(collocated))
    return psi


def velocity_longdouble(tx, fx, fzeta, farea, fmask, collocated: bool):
    tx = np.ascontiguousarray(tx, np.float64)
    u = np.zeros_like(tx)
    psi = np.zeros(tx.shape[0])
    lib().orc_velocity_ld(u, psi, tx, tx.shape[0], np.ascontiguousarray(fx), np.ascontiguousarray(fzeta),
                          np.ascontiguousarray(farea), np.ascontiguousarray(fmask, np.uint8), fx.shape[0],
                          int(collocated))
    return u, psi


def gauss_params(seed: int, nvort: int, amp_scale: float = 1.0):
    c = np.zeros((nvort, 3))
    a = np.zeros(nvort)
    if nvort:
        lib().orc_gauss_params(seed, nvort, amp_scale, c, a)
    return c, a


def ic_eval(x, sbr=0.0, rh_n=0, rh_w=0.0, rh_K=0.0, centers=None, amps=None, sigma=1.0):
    x = np.ascontiguousarray(x, np.float64)
    z = np.zeros(x.shape[0])
    if centers is None:
        centers = np.zeros((1, 3))
        amps = np.zeros(1)
        nv = 0
    else:
        nv = len(amps)
    lib().orc_ic_eval(x, x.shape[0], z, sbr, rh_n, rh_w, rh_K, nv, np.ascontiguousarray(centers),
                      np.ascontiguousarray(amps), sigma)
    return z


class State:
    """The arrays BVESphere hands to BVERK4::advance_timestep."""

    def __init__(self, mesh: Mesh, vzeta, fzeta):
        self.mesh = mesh
        self.vx = mesh.vx.copy()
        self.fx = mesh.fx.copy()
        self.vzeta = np.array(vzeta, np.float64)
        self.fzeta = np.array(fzeta, np.float64)
        self.vvel = np.zeros_like(self.vx)
        self.fvel = np.zeros_like(self.fx)
        self.absv_v = None
        self.absv_f = None

    def init_velocity(self, scheme: int = 0):
        m = self.mesh
        if scheme == 0:
            self.vvel = velocity(self.vx, self.fx, self.fzeta, m.farea, m.fmask, False)
            self.fvel = velocity(self.fx, self.fx, self.fzeta, m.farea, m.fmask, True)
        else:
            self.vvel = velocity_chord(self.vx, self.fx, self.fzeta, m.farea, m.fmask, False)
            self.fvel = velocity_chord(self.fx, self.fx, self.fzeta, m.farea, m.fmask, True)

    def set_lagrangian(self, omega: float):
        self.absv_v = self.vzeta + 2.0 * omega * self.vx[:, 2]
        self.absv_f = self.fzeta + 2.0 * omega * self.fx[:, 2]

    def step(self, dt: float, omega: float, scheme: int = 0):
        """BVERK4::advance_timestep (src/LpmBVERK4_Impl.hpp:42-130)."""
        m = self.mesh
        av = self.absv_v.ctypes.data if scheme else None
        af = self.absv_f.ctypes.data if scheme else None
        lib().orc_rk4_step(scheme, dt, omega, self.vx.shape[0], self.vx, self.vzeta, self.vvel, self.fx.shape[0],
                           self.fx, self.fzeta, self.fvel, m.farea, m.fmask, av, af)


def err_norms_vec(appx, exact, wt):
    out = np.zeros(3)
    lib().orc_err_norms_vec(np.ascontiguousarray(appx), np.ascontiguousarray(exact), np.ascontiguousarray(wt),
                            appx.shape[0], out)
    return out


def err_norms_scalar(appx, exact, wt):
    out = np.zeros(3)
    lib().orc_err_norms_scalar(np.ascontiguousarray(appx), np.ascontiguousarray(exact), np.ascontiguousarray(wt),
                               appx.shape[0], out)
    return out


def sphharm54(x):
    x = np.ascontiguousarray(x, np.float64)
    f = np.zeros(x.shape[0])
    u = np.zeros_like(x)
    lib().orc_sphharm54(x, x.shape[0], f, u)
    return f, u
