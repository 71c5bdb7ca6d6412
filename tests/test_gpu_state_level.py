"""States built on the device from the cached level mesh (lpm_bve_set_state_level):
the same particles as treeInit, the same field as the host evaluation to
rounding, the same dynamics as a host-built state."""
import math

import numpy as np
import pytest

import lpmkokkos_amd as L
from lpmkokkos_amd import _lib
from lpmkokkos_amd import ensemble as E

from helpers import normwise

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def field():
    ic = L.VorticityInitialCondition(sbr=0.5, rh_n=4, rh_w=0.1, rh_K=0.1)
    return ic.with_vortices(5, 12, 0.3, 0.15)


@pytest.mark.parametrize("level", [0, 3, 5])
def test_state_level_matches_tree_init_and_host_field(level):
    m = L.TreeMesh(level)
    ic = field()
    c = L.Context(0)
    c.set_state_level(level, ic)
    assert c.sizes()[:2] == (m.nverts, m.nfaces)
    s = c.get_state()
    # positions: the cached mesh is treeInit's, bit for bit
    assert np.array_equal(s["vx"], m.vx) and np.array_equal(s["fx"], m.fx)
    # field: device FP64 against the host evaluation, to rounding
    hv, hf = ic.eval_many(m.vx), ic.eval_many(m.fx)
    lv = m.fmask == 0
    scale = max(np.abs(hv).max(), np.abs(hf).max())
    assert np.abs(s["vzeta"] - hv).max() < 1e-14 * scale
    assert np.abs(s["fzeta"][lv] - hf[lv]).max() < 1e-14 * scale
    # divided faces are read back from the host evaluation exactly
    assert np.array_equal(s["fzeta"][~lv], hf[~lv])
    c.close()


def test_state_level_mean_removed():
    level = 4
    m = L.TreeMesh(level)
    ic = field()
    a, b = L.Context(0), L.Context(0)
    a.set_state_level(level, ic)
    b.set_state_level(level, ic, subtract_mean=True)
    sa, sb = a.get_state(), b.get_state()
    lv = m.fmask == 0
    mean = L.area_mean(ic.eval_many(m.fx)[lv], m.farea[lv], np.zeros(int(lv.sum()), np.uint8))
    d = sa["fzeta"][lv] - sb["fzeta"][lv]
    assert np.abs(d - mean).max() < 1e-14 * max(1.0, abs(mean))
    assert np.abs(sa["vzeta"] - sb["vzeta"] - mean).max() < 1e-14 * max(1.0, abs(mean))
    # divided faces carry the same shift
    assert np.abs(sa["fzeta"][~lv] - sb["fzeta"][~lv] - mean).max() < 1e-14 * max(1.0, abs(mean))
    total = float(np.sum(np.abs(sb["fzeta"][lv]) * m.farea[lv]))
    assert abs(b.circulation()) < 1e-14 * total
    a.close()
    b.close()


@pytest.mark.parametrize("scheme,prec,tol", [(L.SCHEME_REF, L.PREC_FP64, 1e-12), (L.SCHEME_LAGR, L.PREC_FP64, 1e-12),
                                             (L.SCHEME_LAGR, L.PREC_FP32C, 1e-6)])
def test_state_level_steps_like_host_state(scheme, prec, tol):
    """Five steps from the device-built state against five from the host-built
    state of the same mesh and field (differences start at the field's rounding)."""
    level = 4
    m = L.TreeMesh(level)
    ic = field()
    a = L.Context(0)
    a.set_state_level(level, ic, subtract_mean=True)
    lv = m.fmask == 0
    hv, hf = ic.eval_many(m.vx), ic.eval_many(m.fx)
    mean = L.area_mean(hf[lv], m.farea[lv], np.zeros(int(lv.sum()), np.uint8))
    b = L.Context(0)
    b.set_state(m.vx, hv - mean, m.fx, hf - mean, m.farea, m.fmask)
    for c in (a, b):
        c.advance(0.005, TWO_PI, scheme, prec, nsteps=5)
    sa, sb = a.get_state(), b.get_state()
    assert normwise(sa["vx"], sb["vx"]) < tol and normwise(sa["fx"][lv], sb["fx"][lv]) < tol
    assert normwise(sa["vu"], sb["vu"]) < tol and normwise(sa["fu"][lv], sb["fu"][lv]) < tol
    a.close()
    b.close()


def test_mesh_cache_shared_and_released():
    level = 3
    ic = field()
    a, b = L.Context(0), L.Context(0)
    a.set_state_level(level, ic)
    b.set_state_level(level, ic)
    kept = _lib.C.c_int(-1)
    _lib.call("lpm_mesh_cache_clear", -1, _lib.C.byref(kept))
    assert kept.value >= 1  # in use: kept
    a.close()
    b.advance(0.005, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=2)  # the shared arrays are still there
    s = b.get_state()
    assert np.all(np.isfinite(s["fx"])) and np.all(np.isfinite(s["fu"]))
    b.close()
    _lib.call("lpm_mesh_cache_clear", -1, _lib.C.byref(kept))
    assert kept.value == 0
    # rebuilt on next use
    c = L.Context(0)
    c.set_state_level(level, ic)
    assert np.array_equal(c.get_state()["fx"], L.TreeMesh(level).fx)
    c.close()


def test_member_built_on_device_matches_host_built_member():
    spec = E.MemberSpec(level=4, k_factor=1.02, rh_n=3, nvort=8, amp=0.05, sigma=0.15, vortex_seed=77)
    a = E.make_member(spec)
    b = E.make_member(spec, on_device=False)
    for c in (a, b):
        c.advance(0.005, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=3)
    sa, sb = a.get_state(), b.get_state()
    lv = L.TreeMesh(4).fmask == 0
    assert normwise(sa["vx"], sb["vx"]) < 1e-6 and normwise(sa["fx"][lv], sb["fx"]) < 1e-6
    assert normwise(sa["fzeta"][lv], sb["fzeta"]) < 1e-6
    a.close()
    b.close()


def test_state_level_errors():
    c = L.Context(0)
    with pytest.raises(L.LpmError):
        c.set_state_level(11, field())
    bad = field()
    bad.sigma = 0.0
    with pytest.raises(L.LpmError):
        c.set_state_level(2, bad)
    c.close()


@pytest.mark.parametrize("level", [0, 3, 5])
def test_cubed_state_level_matches_tree_init_and_host_field(level):
    """lpm_bve_set_state_seed_level for CubedSphereSeed: the cached cubed-sphere
    tree is lpm_seed_tree_init's bit for bit, the field the host evaluation's
    to rounding, and the icosahedral tree of the same level is cached apart."""
    m = L.TreeMesh(level, L.MeshSeed(L.CubedSphereSeed))
    ic = field()
    c = L.Context(0)
    c.set_state_level(level, ic, seed_kind=L.SEED_CUBED_SPHERE)
    assert c.sizes()[:2] == (m.nverts, m.nfaces)
    s = c.get_state()
    assert np.array_equal(s["vx"], m.vx) and np.array_equal(s["fx"], m.fx)
    hv, hf = ic.eval_many(m.vx), ic.eval_many(m.fx)
    lv = m.fmask == 0
    scale = max(np.abs(hv).max(), np.abs(hf).max())
    assert np.abs(s["vzeta"] - hv).max() < 1e-14 * scale
    assert np.abs(s["fzeta"][lv] - hf[lv]).max() < 1e-14 * scale
    d = L.Context(0)
    d.set_state_level(level, ic)
    assert d.sizes()[:2] == (L.TreeMesh(level).nverts, L.TreeMesh(level).nfaces)
    c.close()
    d.close()


def test_cubed_state_level_steps_like_host_state():
    """Five FP64 Lagrangian steps from the device-built level-4 cubed-sphere
    state against five from the host-built one."""
    level = 4
    m = L.TreeMesh(level, L.MeshSeed(L.CubedSphereSeed))
    ic = field()
    a = L.Context(0)
    a.set_state_level(level, ic, subtract_mean=True, seed_kind=L.SEED_CUBED_SPHERE)
    lv = m.fmask == 0
    hv, hf = ic.eval_many(m.vx), ic.eval_many(m.fx)
    mean = L.area_mean(hf[lv], m.farea[lv], np.zeros(int(lv.sum()), np.uint8))
    b = L.Context(0)
    b.set_state(m.vx, hv - mean, m.fx, hf - mean, m.farea, m.fmask)
    for c in (a, b):
        c.advance(0.005, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=5)
    sa, sb = a.get_state(), b.get_state()
    assert normwise(sa["vx"], sb["vx"]) < 1e-12 and normwise(sa["fx"][lv], sb["fx"][lv]) < 1e-12
    assert normwise(sa["vu"], sb["vu"]) < 1e-12 and normwise(sa["fu"][lv], sb["fu"][lv]) < 1e-12
    a.close()
    b.close()


def test_state_seed_level_errors():
    c = L.Context(0)
    p = _lib.IcParams(0.0, 4, 0.1, 0.1, 0, None, None, 1.0, 0)
    with pytest.raises(L.LpmError):
        _lib.call("lpm_bve_set_state_seed_level", c.h, 7, 3, _lib.C.byref(p))
    with pytest.raises(L.LpmError):
        c.set_state_level(11, field(), seed_kind=L.SEED_CUBED_SPHERE)
    c.close()
