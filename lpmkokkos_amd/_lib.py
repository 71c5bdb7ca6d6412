"""ctypes binding of liblpm_bve.so (the C ABI declared in include/lpm_bve.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (make in
lpmkokkos_amd/csrc).  There is no fallback: if the library is missing,
loading raises, so nothing here ever computes on the CPU in its place.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "liblpm_bve.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "lpm_bve.h")

LPM_OK = 0
PREC_FP64, PREC_FP32C = 0, 1
SCHEME_REF, SCHEME_LAGR = 0, 1
SEED_ICOS_TRI, SEED_CUBED_SPHERE = 0, 1
VTK_LEGACY_ASCII, VTK_LEGACY_BINARY, VTK_XML_ASCII, VTK_XML_BINARY = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_dpn = C.c_void_p  # nullable double*
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")



class IcParams(C.Structure):
    """lpm_ic_params of include/lpm_bve.h."""
    _fields_ = [("sbr", C.c_double), ("rh_n", C.c_int), ("rh_w", C.c_double), ("rh_K", C.c_double),
                ("nvort", C.c_int), ("centers", C.c_void_p), ("amps", C.c_void_p), ("sigma", C.c_double),
                ("subtract_mean", C.c_int)]


class VtkArray(C.Structure):
    """lpm_vtk_array of include/lpm_bve.h."""
    _fields_ = [("name", C.c_char_p), ("ncomp", C.c_int), ("on_faces", C.c_int), ("data", C.c_void_p)]


# (name, argtypes) for every entry point of include/lpm_bve.h
SIGNATURES = {
    "lpm_last_error": [],
    "lpm_version": [],
    "lpm_device_count": [C.POINTER(C.c_int)],
    "lpm_icos_alloc": [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "lpm_icos_tree_init": [C.c_int, C.c_char_p, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn],
    "lpm_seed_alloc": [C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "lpm_seed_tree_init": [C.c_int, C.c_int, C.c_char_p, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn],
    "lpm_seed_mesh_write_nc": [C.c_int, C.c_int, C.c_char_p, C.c_char_p],
    "lpm_seed_mesh_write_vtk": [C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int],
    "lpm_seed_mesh_write_vtk_arrays": [C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(VtkArray), C.c_int],
    "lpm_ic_gauss_params": [C.c_uint64, C.c_int, C.c_double, _dp, _dp],
    "lpm_ic_eval": [_dp, C.c_int64, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _dpn, _dpn,
                    C.c_double],
    "lpm_area_mean": [_dp, _dp, _bp, C.c_int64, C.POINTER(C.c_double)],
    "lpm_err_norms_vec": [_dpn, _dp, _dp, _dp, C.c_int64, _dp],
    "lpm_err_norms_scalar": [_dp, _dp, _dp, C.c_int64, _dp],
    "lpm_velocity_tangent": [_dp, _dp, C.c_int64, _dp],
    "lpm_poisson_y54": [_dp, C.c_int64, _dpn, _dpn],
    "lpm_mesh_write_nc": [C.c_int, C.c_char_p, C.c_char_p],
    "lpm_mesh_write_vtk": [C.c_int, C.c_char_p, C.c_char_p, C.c_int],
    "lpm_mesh_read_nc": [C.c_char_p, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn],
    "lpm_rng_draws": [C.c_uint64, C.c_int64, _dpn, C.c_int64, _dpn],
    "lpm_bve_vertex_velocity": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64, C.c_void_p],
    "lpm_bve_face_velocity": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "lpm_bve_vertex_streamfn": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int64, C.c_void_p],
    "lpm_bve_face_streamfn": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "lpm_bve_vertex_solve": [_dpn, _dpn, _dpn, C.c_int64, _dpn, _dpn, _dpn, _dpn, C.c_int64, C.c_void_p],
    "lpm_bve_face_solve": [_dpn, _dpn, _dpn, _dpn, _dpn, _dpn, C.c_int64, C.c_void_p],
    "lpm_bve_functor_static": [_dpn, _dpn, _dpn, C.c_int64, _dpn, _dpn, _dpn, _dpn, C.c_int64, C.c_int, C.c_void_p],
    "lpm_bve_create": [C.c_int, C.POINTER(C.c_void_p)],
    "lpm_bve_create_on_stream": [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)],
    "lpm_bve_destroy": [C.c_void_p],
    "lpm_bve_set_state": [C.c_void_p, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, _bp],
    "lpm_bve_set_state_level": [C.c_void_p, C.c_int, C.POINTER(IcParams)],
    "lpm_bve_set_state_seed_level": [C.c_void_p, C.c_int, C.c_int, C.POINTER(IcParams)],
    "lpm_mesh_cache_clear": [C.c_int, C.POINTER(C.c_int)],
    "lpm_bve_set_velocity": [C.c_void_p, _dp, _dp],
    "lpm_bve_solve": [C.c_void_p, C.c_int, C.c_int, C.c_int],
    "lpm_bve_streamfn": [C.c_void_p, C.c_int, C.c_int],
    "lpm_bve_advance": [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int],
    "lpm_bve_get_state": [C.c_void_p, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn, _dpn],
    "lpm_bve_checkpoint": [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_double],
    "lpm_bve_restore": [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p],
    "lpm_bve_write_vtk": [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_int],
    "lpm_bve_write_vtk_arrays": [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_int],
    "lpm_bve_circulation": [C.c_void_p, C.POINTER(C.c_double)],
    "lpm_bve_synchronize": [C.c_void_p],
    "lpm_bve_query": [C.c_void_p, C.POINTER(C.c_int)],
    "lpm_bve_sizes": [C.c_void_p, _i64],
    "lpm_bve_set_launch": [C.c_void_p, C.c_int, C.c_int],
    "lpm_bve_set_near_factor": [C.c_void_p, C.c_double],
    "lpm_bve_set_source_staging": [C.c_void_p, C.c_int],
    "lpm_bve_stream": [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int)],
    "lpm_bve_last_pair_timing": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "lpm_bve_enable_timing": [C.c_void_p, C.c_int],
    "lpm_bve_state_bytes": [C.c_int64, C.c_int64, C.POINTER(C.c_int64)],
    "lpm_shape_rule_set": [C.c_int64, C.c_int, C.c_int],
    "lpm_shape_rule_get": [C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "lpm_shape_rules_clear": [],
    "lpm_bve_autotune": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                         C.POINTER(C.c_double)],
    "lpm_ens_create": [C.c_int, C.POINTER(C.c_void_p)],
    "lpm_ens_destroy": [C.c_void_p],
    "lpm_ens_add": [C.c_void_p, C.c_void_p],
    "lpm_ens_remove": [C.c_void_p, C.c_void_p],
    "lpm_ens_size": [C.c_void_p, C.POINTER(C.c_int)],
    "lpm_ens_set_blocks_per_cu": [C.c_void_p, C.c_int],
    "lpm_ens_autotune": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                         C.c_void_p, C.c_int],
    "lpm_ens_rules_clear": [],
    "lpm_ens_autotune_shapes": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int)],
    "lpm_ens_advance": [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int],
    "lpm_ens_synchronize": [C.c_void_p],
    "lpm_ens_query": [C.c_void_p, C.POINTER(C.c_int)],
    "lpm_ens_stream": [C.c_void_p, C.POINTER(C.c_void_p)],
}

_lib = None


class LpmError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load liblpm_bve.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C lpmkokkos_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.lpm_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(fn: str, rc: int) -> None:
    if rc != LPM_OK:
        raise LpmError(fn, rc, lib().lpm_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(name, getattr(lib(), name)(*args))


def ptr(a) -> int | None:
    """Data pointer of a numpy array or torch tensor (None passes NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()
