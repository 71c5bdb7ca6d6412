"""The allocation-free functor entry point (lpm_bve_functor_static): its
scratch is a fixed static device array, so it calls no hipMalloc or hipFree.
Checked against the allocating functor entry points it mirrors
(lpm_bve_vertex_solve, lpm_bve_face_solve), bit for bit, on two calls in a
row, the second with a smaller input; sizes beyond its fixed scratch are
refused before any device work."""
import numpy as np
import pytest

import lpmkokkos_amd as L
from helpers import DeviceArray, rh_field

pytestmark = pytest.mark.gpu

P = L._lib.ptr


def run_pair(m, fz):
    """(allocating, static) results: vertex solve and face solve."""
    d = {k: DeviceArray(v) for k, v in (("vx", m.vx), ("fx", m.fx), ("z", fz), ("a", m.farea), ("m", m.fmask))}
    out = {}
    for which in ("alloc", "static"):
        pv, uv = DeviceArray.zeros(m.nv), DeviceArray.zeros(m.vx.shape)
        pf, uf = DeviceArray.zeros(m.nf), DeviceArray.zeros(m.fx.shape)
        if which == "alloc":
            L._lib.call("lpm_bve_vertex_solve", P(pv), P(uv), P(d["vx"]), m.nv, P(d["fx"]), P(d["z"]), P(d["a"]),
                        P(d["m"]), m.nf, None)
            L._lib.call("lpm_bve_face_solve", P(pf), P(uf), P(d["fx"]), P(d["z"]), P(d["a"]), P(d["m"]), m.nf, None)
        else:
            L._lib.call("lpm_bve_functor_static", P(pv), P(uv), P(d["vx"]), m.nv, P(d["fx"]), P(d["z"]), P(d["a"]),
                        P(d["m"]), m.nf, 0, None)
            L._lib.call("lpm_bve_functor_static", P(pf), P(uf), P(d["fx"]), m.nf, P(d["fx"]), P(d["z"]), P(d["a"]),
                        P(d["m"]), m.nf, 1, None)
        out[which] = [x.numpy() for x in (pv, uv, pf, uf)]
        for x in (pv, uv, pf, uf):
            x.free()
    # stream function only (no velocity output: the velocity goes to scratch)
    ps = DeviceArray.zeros(m.nf)
    L._lib.call("lpm_bve_functor_static", P(ps), None, P(d["fx"]), m.nf, P(d["fx"]), P(d["z"]), P(d["a"]), P(d["m"]),
                m.nf, 1, None)
    out["static_psi_only"] = ps.numpy()
    ps.free()
    for x in d.values():
        x.free()
    return out


def test_static_scratch_matches_allocating_functors(omesh):
    lv_prev = None
    for level in (5, 3):  # the second call is smaller and reuses the same scratch
        m = omesh(level)
        fz = rh_field(m.fx) + 0.3
        out = run_pair(m, fz)
        lv = m.fmask == 0
        a, s = out["alloc"], out["static"]
        assert np.array_equal(a[0], s[0]) and np.array_equal(a[1], s[1])
        assert np.array_equal(a[2][lv], s[2][lv]) and np.array_equal(a[3][lv], s[3][lv])
        assert np.array_equal(out["static_psi_only"][lv], a[2][lv])
        assert np.abs(a[1]).max() > 0 and np.all(np.isfinite(s[1]))
        lv_prev = lv
    assert lv_prev is not None


def test_static_scratch_refuses_oversized():
    m = L.TreeMesh(7)  # 436,900 faces: the face evaluation needs more than the fixed scratch
    with pytest.raises(L.LpmError) as e:
        L._lib.call("lpm_bve_functor_static", 1, 1, 1, m.nfaces, 1, 1, 1, 1, m.nfaces, 1, None)
    assert "scratch" in str(e.value)
