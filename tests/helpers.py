"""Shared builders for the parity tests (oracle side is the checker)."""
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
