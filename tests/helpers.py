"""Shared builders for the parity tests (the oracle supplies the expected values)."""
import numpy as np

from oracle import oracle as O


class LeafMesh:
    """An oracle mesh restricted to vertices and active faces.  Summing only
    the active faces gives the same bits as the reference's masked loop over
    all tree faces (masked faces add an exact zero); the test
    test_leaf_restriction_is_exact pins that."""

    def __init__(self, m):
        lv = m.fmask == 0
        self.full = m
        self.vx = m.vx
        self.fx = np.ascontiguousarray(m.fx[lv])
        self.farea = np.ascontiguousarray(m.farea[lv])
        self.fmask = np.zeros(int(lv.sum()), np.uint8)
        self.nv = m.nv
        self.nleaves = int(lv.sum())


def rh_field(x, n=4, w=0.1, K=0.1):
    return O.ic_eval(x, rh_n=n, rh_w=w, rh_K=K)


def gauss_field(m, seed, nvort=32, sigma=0.1, amp=1.0, rh=None):
    """Sum of seeded Gaussian vortices (plus an optional RH wave), minus the
    area-weighted mean over active faces."""
    c, a = O.gauss_params(seed, nvort, amp)
    kw = dict(centers=c, amps=a, sigma=sigma)
    if rh:
        kw.update(rh_n=rh[0], rh_w=rh[1], rh_K=rh[2])
    vz = O.ic_eval(m.vx, **kw)
    fz = O.ic_eval(m.fx, **kw)
    lv = m.fmask == 0
    mu = float(np.sum(fz[lv] * m.farea[lv]) / np.sum(m.farea[lv]))
    return vz - mu, fz - mu


def normwise(a, b):
    """max |a - b| / max |b| over rows (vector L-inf of src/LpmErrorNorms.cpp:33-44)."""
    a = np.asarray(a).reshape(len(b), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float(np.linalg.norm(a - b, axis=1).max() / np.linalg.norm(b, axis=1).max())


class DeviceArray:
    """Device buffer allocated through the HIP runtime liblpm_bve.so already
    loaded (one HIP runtime per process), for the functor entry points."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C

            import lpmkokkos_amd
            lpmkokkos_amd.lib()  # loads libamdhip64 as the library's dependency
            h = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host: np.ndarray):
        import ctypes as C
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(p), max(self.nbytes, 8)) == 0
        self.ptr = p.value
        h = np.ascontiguousarray(host)
        assert self.hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, 1) == 0  # host to device

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        return cls(np.zeros(shape, dtype))

    def data_ptr(self):
        return self.ptr

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
