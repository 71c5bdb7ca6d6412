"""Parity of the HIP path (through the C ABI of liblpm_bve.so) with the CPU
oracle.  Tolerances (normwise, src/LpmErrorNorms.cpp:33-44):
  FP64 reference scheme: velocity 1e-11, positions 1e-9 after 100 steps;
  single evaluations 1e-13 (summation order differs from the oracle's);
  FP32C vs FP64: velocity 1e-6, stream function 1e-6.
"""
import math

import numpy as np
import pytest

import lpmkokkos_amd as L
from helpers import LeafMesh, gauss_field, normwise, rh_field
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi


def make_ctx(lm, vz, fz):
    c = L.Context(0)
    c.set_state(lm.vx, vz, lm.fx, fz, lm.farea, lm.fmask)
    return c


def test_functor_entry_points(omesh):
    """BVE{Vertex,Face}{Velocity,StreamFn,Solve} on caller-owned device arrays over
    all tree faces (masked loop), against the oracle at vertices and active faces."""
    from helpers import DeviceArray
    m = omesh(4)
    fz = rh_field(m.fx)
    vx, fx, z, a, msk = (DeviceArray(h) for h in (m.vx, m.fx, fz, m.farea, m.fmask))
    uv = DeviceArray.zeros(m.vx.shape)
    uf = DeviceArray.zeros(m.fx.shape)
    pv = DeviceArray.zeros(m.nv)
    pf = DeviceArray.zeros(m.nf)
    P = L._lib.ptr
    L._lib.call("lpm_bve_vertex_velocity", P(uv), P(vx), m.nv, P(fx), P(z), P(a), P(msk), m.nf, None)
    L._lib.call("lpm_bve_face_velocity", P(uf), P(fx), P(z), P(a), P(msk), m.nf, None)
    L._lib.call("lpm_bve_vertex_streamfn", P(pv), P(vx), m.nv, P(fx), P(z), P(a), P(msk), m.nf, None)
    L._lib.call("lpm_bve_face_streamfn", P(pf), P(fx), P(z), P(a), P(msk), m.nf, None)
    lv = m.fmask == 0
    ov = O.velocity(m.vx, m.fx, fz, m.farea, m.fmask, False)
    of = O.velocity(m.fx, m.fx, fz, m.farea, m.fmask, True)
    opv = O.streamfn(m.vx, m.fx, fz, m.farea, m.fmask, False)
    opf = O.streamfn(m.fx, m.fx, fz, m.farea, m.fmask, True)
    assert normwise(uv.numpy(), ov) < 1e-13
    assert normwise(uf.numpy()[lv], of[lv]) < 1e-13
    assert normwise(pv.numpy(), opv) < 1e-13
    assert normwise(pf.numpy()[lv], opf[lv]) < 1e-13
    # BVEVertexSolve / BVEFaceSolve: both fields in one pass, the same sums
    uv2, uf2 = DeviceArray.zeros(m.vx.shape), DeviceArray.zeros(m.fx.shape)
    pv2, pf2 = DeviceArray.zeros(m.nv), DeviceArray.zeros(m.nf)
    L._lib.call("lpm_bve_vertex_solve", P(pv2), P(uv2), P(vx), m.nv, P(fx), P(z), P(a), P(msk), m.nf, None)
    L._lib.call("lpm_bve_face_solve", P(pf2), P(uf2), P(fx), P(z), P(a), P(msk), m.nf, None)
    assert np.array_equal(uv2.numpy(), uv.numpy()) and np.array_equal(pv2.numpy(), pv.numpy())
    assert np.array_equal(uf2.numpy()[lv], uf.numpy()[lv]) and np.array_equal(pf2.numpy()[lv], pf.numpy()[lv])
    with pytest.raises(L.LpmError):
        L._lib.call("lpm_bve_face_solve", None, P(uf2), P(fx), P(z), P(a), P(msk), m.nf, None)
    for d in (vx, fx, z, a, msk, uv, uf, pv, pf, uv2, uf2, pv2, pf2):
        d.free()


def test_coincident_target_leaves_other_targets_alone(omesh):
    """A vertex target exactly on a source that is not its own (1 - x.y = 0 exactly, as a
    divided face's centre on its middle child's) makes only its own row singular, as in the
    reference's literal division: the FP64 sums share one reciprocal over four pairs of one
    target, never across targets.  Every other target matches the oracle."""
    from helpers import DeviceArray
    m = omesh(3)
    vx, fx = m.vx.copy(), m.fx.copy()
    fz = rh_field(fx)
    j = int(np.flatnonzero(m.fmask == 0)[5])
    fx[j] = (0.0, 0.0, 1.0)
    vx[0] = (0.0, 0.0, 1.0)
    assert 1.0 - float(vx[0] @ fx[j]) == 0.0
    dv, df, dz, da, dm = (DeviceArray(h) for h in (vx, fx, fz, m.farea, m.fmask))
    du = DeviceArray.zeros(vx.shape)
    L._lib.call("lpm_bve_vertex_velocity", L._lib.ptr(du), L._lib.ptr(dv), m.nv, L._lib.ptr(df),
                L._lib.ptr(dz), L._lib.ptr(da), L._lib.ptr(dm), m.nf, None)
    u = du.numpy()
    ov = O.velocity(vx, fx, fz, m.farea, m.fmask, False)
    assert m.nv > 256  # targets i and i + 256 share a lane at this size
    assert not np.all(np.isfinite(u[0])) and not np.all(np.isfinite(ov[0]))
    assert np.all(np.isfinite(u[1:]))
    assert normwise(u[1:], ov[1:]) < 1e-13
    for d in (dv, df, dz, da, dm, du):
        d.free()


@pytest.mark.parametrize("level", [0, 2, 4])
def test_solve_fp64_ref(omesh, level):
    m = omesh(level)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_REF, want_psi=True)
    s = c.get_state()
    ov = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False)
    of = O.velocity(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True)
    assert normwise(s["vu"], ov) < 1e-13
    assert normwise(s["fu"], of) < 1e-13
    assert normwise(s["vpsi"], O.streamfn(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False)) < 1e-13
    assert normwise(s["fpsi"], O.streamfn(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True)) < 1e-13
    c.close()


def test_tree_faces_drop_in(omesh):
    """Passing every tree face (the reference's arrays): active faces and
    vertices match the oracle; divided faces keep position, zero velocity."""
    m = omesh(3)
    vz, fz = rh_field(m.vx), rh_field(m.fx)
    c = L.Context(0)
    c.set_state(m.vx, vz, m.fx, fz, m.farea, m.fmask)
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    s = c.get_state()
    lv = m.fmask == 0
    of = O.velocity(m.fx, m.fx, fz, m.farea, m.fmask, True)
    assert normwise(s["fu"][lv], of[lv]) < 1e-13
    assert np.all(s["fu"][~lv] == 0) and np.array_equal(s["fx"][~lv], m.fx[~lv])
    assert c.sizes() == (m.nv, m.nf, m.nleaves, m.nv + m.nleaves)
    c.close()


def test_level4_rh_wave_100_steps_ref(omesh):
    """Level-4 Rossby-Haurwitz wave (n=4, w=K=0.1), Omega=2 pi, dt=0.01, 100 steps
    of the reference scheme in FP64: velocity <= 1e-11, positions <= 1e-9 (normwise)."""
    m = omesh(4)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    o = O.State(lm, vz, fz)
    o.init_velocity(0)
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    c.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=100)
    for _ in range(100):
        o.step(0.01, TWO_PI, 0)
    s = c.get_state()
    ev = max(normwise(s["vu"], o.vvel), normwise(s["fu"], o.fvel))
    ex = max(normwise(s["vx"], o.vx), normwise(s["fx"], o.fx))
    ez = max(normwise(s["vzeta"], o.vzeta), normwise(s["fzeta"], o.fzeta))
    print(f"level 4, 100 steps: velocity {ev:.3e} positions {ex:.3e} vorticity {ez:.3e}")
    assert ev <= 1e-11 and ex <= 1e-9 and ez <= 1e-9
    c.close()


def test_copy_in_copy_out_binding_matches_resident(omesh):
    """The reference-side binding (every advance_timestep copies the caller's
    arrays in with lpm_bve_set_state + lpm_bve_set_velocity, steps once, and
    copies them out with lpm_bve_get_state; src/LpmBVERK4_Impl.hpp:42-59)
    over the level-4 Rossby-Haurwitz configuration for 100 reference-scheme
    steps equals a resident run of lpm_bve_advance, bit for bit.  The tree
    faces (divided ones included) go in and out as the reference holds them;
    set_state updates the resident buffers in place."""
    import time
    m = omesh(4)
    vz, fz = rh_field(m.vx), rh_field(m.fx)
    res = L.Context(0)
    res.set_state(m.vx, vz, m.fx, fz, m.farea, m.fmask)
    res.solve(L.PREC_FP64, L.SCHEME_REF)
    s = res.get_state()
    vx, vzeta, vu, fx, fzeta, fu = (s[k].copy() for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"))
    res.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=100)
    c = L.Context(0)
    t0 = time.perf_counter()
    for _ in range(100):
        c.set_state(vx, vzeta, fx, fzeta, m.farea, m.fmask)
        c.set_velocity(vu, fu)
        c.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=1)
        s = c.get_state()
        vx, vzeta, vu, fx, fzeta, fu = (s[k] for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"))
    dt = time.perf_counter() - t0
    r = res.get_state()
    for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
        assert np.array_equal(r[k], s[k]), k
    print(f"copy-in/copy-out binding: {1e3 * dt / 100:.2f} ms per step at level 4")
    res.close()
    c.close()


def test_lagr_scheme_fp64(omesh):
    m = omesh(3)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=3, nvort=8, sigma=0.15, rh=(4, 0.1, 0.1))
    fz = fz[m.fmask == 0]
    o = O.State(lm, vz, fz)
    o.set_lagrangian(TWO_PI)
    o.init_velocity(1)
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    c.advance(0.005, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=20)
    for _ in range(20):
        o.step(0.005, TWO_PI, 1)
    s = c.get_state()
    assert max(normwise(s["vu"], o.vvel), normwise(s["fu"], o.fvel)) < 1e-11
    assert max(normwise(s["vx"], o.vx), normwise(s["fx"], o.fx)) < 1e-11
    assert np.abs(np.linalg.norm(s["fx"], axis=1) - 1).max() < 1e-15
    assert np.abs(s["fzeta"] + 2 * TWO_PI * s["fx"][:, 2] - (fz + 2 * TWO_PI * lm.fx[:, 2])).max() < 1e-12
    c.close()


@pytest.mark.parametrize("level", [5, 6])
def test_fp32c_vs_fp64(omesh, level):
    """FP32C pair sums against FP64 on the Gaussian-vortex field plus RH wave:
    velocity and stream function within 1e-6 normwise."""
    m = omesh(level)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1, rh=(4, 0.1, 0.1))
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    a = c.get_state()
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=True)
    b = c.get_state()
    eu = max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"]))
    ep = max(normwise(b["vpsi"], a["vpsi"]), normwise(b["fpsi"], a["fpsi"]))
    print(f"level {level}: FP32C velocity {eu:.3e} psi {ep:.3e}")
    assert eu < 1e-6 and ep < 1e-6
    # the FP64 chord form agrees with the oracle's chord form
    ov = O.velocity_chord(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False)
    assert normwise(a["vu"], ov) < 1e-13
    c.close()


@pytest.mark.parametrize("level", [6, 7])
def test_fp32c_step_kernel_vs_fp64(omesh, level):
    """The velocity-only FP32C kernel (the one inside RK4 steps; its far form
    shares one reciprocal among four pairs) against FP64 for every launch shape
    and both source stagings: within 1e-6 normwise, and within 1e-6 of the
    FP32C kernel that also sums the stream function."""
    m = omesh(level)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1, rh=(4, 0.1, 0.1))
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    a = c.get_state()
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=True)
    p = c.get_state()
    for lds in (True, False):
        for R in (1, 2, 3, 4):
            c.set_source_staging(lds)
            c.set_launch(R, 0)
            c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
            b = c.get_state()
            eu = max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"]))
            ep = max(normwise(b["vu"], p["vu"]), normwise(b["fu"], p["fu"]))
            print(f"level {level} lds {lds} R {R}: velocity vs FP64 {eu:.3e}, vs psi kernel {ep:.3e}")
            assert eu < 1e-6 and ep < 1e-6, (lds, R)
    c.close()


def test_graph_replay_matches_eager(omesh):
    """Steps replayed from the captured hipGraph give the same bits as eager
    launches, replay after replay with eager work in between."""
    m = omesh(4)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    a = make_ctx(lm, vz, fz)
    b = make_ctx(lm, vz, fz)
    for c in (a, b):
        c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    for rep in range(3):
        a.advance(0.01, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=2, use_graph=True)
        b.advance(0.01, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=2, use_graph=False)
        sa, sb = a.get_state(), b.get_state()
        a.circulation()
        for k in ("vx", "fx", "vu", "fu", "fzeta"):
            assert np.array_equal(sa[k], sb[k]), (rep, k)
    a.close()
    b.close()


@pytest.mark.parametrize("R,splits", [(1, 1), (2, 3), (3, 0), (3, 5), (4, 0), (4, 7)])
def test_launch_shapes_agree(omesh, R, splits):
    m = omesh(4)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    c = make_ctx(lm, vz, fz)
    ov = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False)
    of = O.velocity(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True)
    c.set_launch(R, splits)
    for prec, scheme, tol in ((L.PREC_FP64, L.SCHEME_REF, 1e-13), (L.PREC_FP32C, L.SCHEME_LAGR, 1e-6)):
        c.solve(prec, scheme)
        s = c.get_state()
        assert normwise(s["vu"], ov) < tol and normwise(s["fu"], of) < tol, (prec, R, splits)
    c.close()


def test_level7_gaussian_vortices_subset(omesh):
    """Level 7 (163,842 vertices, 327,680 faces), 32 Gaussian vortices (sigma 0.1,
    seed 1, mean removed):
    FP64 reference-scheme velocity against the oracle on every 997th target,
    within the 1e-11 velocity bound."""
    from oracle.oracle import Mesh
    m = Mesh(7)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1)
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    s = c.get_state()
    stride = 997
    ov = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    of = O.velocity(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    # the literal 1 - x.y keeps ~5 digits for nearest pairs at level 7; FMA vs
    # separately rounded dot products differ there by ~1e-11 relative
    ev, ef = normwise(s["vu"][::stride], ov), normwise(s["fu"][::stride], of)
    print(f"level 7 subset: vertex {ev:.3e} face {ef:.3e}")
    assert ev < 1e-11 and ef < 1e-11
    c.close()


def test_level8_lagrangian_subset():
    """Level 8 (655,362 vertices, 1,310,720 faces; the benchmark's workload), 32
    Gaussian vortices (sigma 0.1, seed 1, mean removed): the FP64 Lagrangian-scheme
    velocity against the oracle's chord-form sums on every 4999th target (no
    cancellation in |y - x|^2 / 2, so to 1e-12), the FP32C step kernel against
    the FP64 sums at every target (1e-6), and the reference scheme's FP64 sums
    (literal 1 - x.y) against the oracle's on the same subset (1e-11)."""
    from oracle.oracle import Mesh
    m = Mesh(8)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1)
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    a = c.get_state()
    stride = 4999
    ov = O.velocity_chord(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    of = O.velocity_chord(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    ev, ef = normwise(a["vu"][::stride], ov), normwise(a["fu"][::stride], of)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    b = c.get_state()
    e32 = max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"]))
    # the reference scheme's literal 1 - x.y, within the 1e-11 velocity bound
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    r = c.get_state()
    rv = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    rf = O.velocity(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    erv, erf = normwise(r["vu"][::stride], rv), normwise(r["fu"][::stride], rf)
    print(f"level 8 subset: FP64 chord vertex {ev:.3e} face {ef:.3e}; FP32C vs FP64 {e32:.3e}; "
          f"FP64 dot vertex {erv:.3e} face {erf:.3e}")
    assert ev < 1e-12 and ef < 1e-12 and e32 < 1e-6 and erv < 1e-11 and erf < 1e-11
    c.close()


@pytest.mark.timeout(300)
def test_level9_subset():
    """The largest size run here: level 9 (2,621,442 vertices, 5,242,880 faces,
    4.1e13 pairs per evaluation), where the bound on the split partials, not the
    workgroup length, sets the default launch (18 source splits).  The same 32
    Gaussian vortices: FP64 Lagrangian-scheme velocity against the oracle's
    chord-form sums on every 19997th target (1e-12), FP32C against FP64 at
    every target (1e-6)."""
    from oracle.oracle import Mesh
    m = Mesh(9)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1)
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    print(f"level 9: {lm.vx.shape[0]} vertices, {fz.size} faces", flush=True)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    a = c.get_state()
    stride = 19997
    ov = O.velocity_chord(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    of = O.velocity_chord(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    ev, ef = normwise(a["vu"][::stride], ov), normwise(a["fu"][::stride], of)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    b = c.get_state()
    e32 = max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"]))
    print(f"level 9 subset: FP64 chord vertex {ev:.3e} face {ef:.3e}; FP32C vs FP64 {e32:.3e}")
    assert ev < 1e-12 and ef < 1e-12 and e32 < 1e-6
    c.close()


def test_cubed_sphere_level8_subset():
    """The cubed-sphere seed at level 8 (393,218 vertices, 393,216 quadrilateral
    leaves), 32 Gaussian vortices: FP64 Lagrangian-scheme velocity against the
    oracle's chord-form sums on every 997th target (1e-12), FP32C against FP64
    at every target (1e-6), the reference scheme's literal 1 - x.y against the
    oracle on the same subset (1e-11)."""
    m = O.Mesh(8, kind="cubed")
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=1, nvort=32, sigma=0.1)
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    a = c.get_state()
    stride = 997
    ov = O.velocity_chord(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    of = O.velocity_chord(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    ev, ef = normwise(a["vu"][::stride], ov), normwise(a["fu"][::stride], of)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    b = c.get_state()
    e32 = max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"]))
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    r = c.get_state()
    rv = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False, stride=stride)[::stride]
    rf = O.velocity(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True, stride=stride)[::stride]
    erv, erf = normwise(r["vu"][::stride], rv), normwise(r["fu"][::stride], rf)
    print(f"cubed level 8 subset: FP64 chord vertex {ev:.3e} face {ef:.3e}; FP32C vs FP64 {e32:.3e}; "
          f"FP64 dot vertex {erv:.3e} face {erf:.3e}")
    assert ev < 1e-12 and ef < 1e-12 and e32 < 1e-6 and erv < 1e-11 and erf < 1e-11
    c.close()


def test_solid_body_rotation_gpu(omesh):
    """tests/LpmBVETest.cpp: zeta = 2 Omega z moves particles rigidly.  The
    renormalised scheme tracks the exact rotation at level 4; the reference
    scheme matches the oracle at level 3.  (At level 4, dt = 0.01, the
    reference scheme's unnormalised stage positions sit ~1e-3 off the sphere,
    so 1 - x.y <= 0 for nearest neighbours and both the oracle and this path
    diverge -- see test_reference_scheme_sbr_instability.)"""
    m = omesh(4)
    lm = LeafMesh(m)
    vz = O.ic_eval(lm.vx, sbr=1.0)
    fz = O.ic_eval(lm.fx, sbr=1.0)
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR)
    dt, n = 0.01, 20
    c.advance(dt, TWO_PI, L.SCHEME_LAGR, L.PREC_FP64, nsteps=n)
    s = c.get_state()
    t = dt * n
    cs, sn = math.cos(TWO_PI * t), math.sin(TWO_PI * t)
    ex = np.stack([lm.vx[:, 0] * cs - lm.vx[:, 1] * sn, lm.vx[:, 1] * cs + lm.vx[:, 0] * sn, lm.vx[:, 2]], 1)
    assert np.abs(s["vx"] - ex).max() < 1e-2
    assert np.abs(np.einsum("ij,ij->i", s["vu"], s["vx"])).max() < 1e-12
    c.close()
    m3 = omesh(3)
    lm3 = LeafMesh(m3)
    vz3, fz3 = O.ic_eval(lm3.vx, sbr=1.0), O.ic_eval(lm3.fx, sbr=1.0)
    c = make_ctx(lm3, vz3, fz3)
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    c.advance(dt, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=n)
    o = O.State(lm3, vz3, fz3)
    o.init_velocity(0)
    for _ in range(n):
        o.step(dt, TWO_PI, 0)
    s = c.get_state()
    assert normwise(s["vx"], o.vx) < 1e-12 and normwise(s["fu"], o.fvel) < 1e-11
    c.close()


def test_circulation(omesh):
    m = omesh(5)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    c = make_ctx(lm, vz, fz)
    g0 = c.circulation()
    assert abs(g0 - float(np.sum(fz * lm.farea))) < 1e-12
    # zeta = absv - 2 Omega z: the sum moves only with the quadrature of z dA
    c.advance(0.01, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=5)
    assert abs(c.circulation() - g0) < 1e-6 * float(np.sum(np.abs(fz) * lm.farea))
    c.close()


@pytest.mark.parametrize("lds,near", [(False, 4.0), (True, 4.0), (True, 2.0), (True, 1e30), (False, 1e30)])
def test_fp32c_variants_agree(omesh, lds, near):
    """Source staging (scalar cache or LDS) and the near/far threshold change
    only rounding: every variant stays within the FP32C bound of FP64."""
    m = omesh(5)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=2, nvort=16, sigma=0.1)
    fz = fz[m.fmask == 0]
    c = make_ctx(lm, vz, fz)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    a = c.get_state()
    c.set_source_staging(lds)
    c.set_near_factor(near)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=True)
    b = c.get_state()
    assert max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"])) < 1e-6
    assert max(normwise(b["vpsi"], a["vpsi"]), normwise(b["fpsi"], a["fpsi"])) < 1e-6
    c.close()


@pytest.mark.parametrize("drop,near", [(0, 2.0), (137, 2.0), (137, 4.0)])
def test_fp32c_padded_tiles(omesh, drop, near):
    """Source counts that leave the last tile partly padded (level 2: 320
    sources; level 5 with `drop` faces removed): padding records add nothing
    and never spoil the far form's shared reciprocals, in every launch shape."""
    m = omesh(2 if drop == 0 else 5)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=3, nvort=16, sigma=0.2)
    fz = fz[m.fmask == 0]
    keep = np.ones(len(fz), bool)
    keep[np.random.default_rng(5).choice(len(fz), drop, replace=False)] = False
    assert keep.sum() % 256 != 0
    c = L.Context(0)
    c.set_state(lm.vx, vz, lm.fx[keep], fz[keep], lm.farea[keep], lm.fmask[keep])
    c.set_near_factor(near)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    a = c.get_state()
    for R in (1, 2, 3, 4):
        c.set_launch(R, 0)
        for psi in (False, True):
            c.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=psi)
            b = c.get_state()
            assert np.all(np.isfinite(b["vu"])) and np.all(np.isfinite(b["fu"]))
            assert max(normwise(b["vu"], a["vu"]), normwise(b["fu"], a["fu"])) < 1e-6, (R, psi)
            if psi:
                assert max(normwise(b["vpsi"], a["vpsi"]), normwise(b["fpsi"], a["fpsi"])) < 1e-6, R
    c.close()


@pytest.mark.parametrize("lds", [True, False])
def test_fp64_staging_variants(omesh, lds):
    """FP64 pair sums read through LDS or the scalar cache agree with the oracle."""
    m = omesh(4)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    c = make_ctx(lm, vz, fz)
    c.set_source_staging(lds)
    for scheme, fn in ((L.SCHEME_REF, O.velocity), (L.SCHEME_LAGR, O.velocity_chord)):
        c.solve(L.PREC_FP64, scheme, want_psi=True)
        s = c.get_state()
        ov = O.velocity(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False) if scheme == L.SCHEME_REF else \
            O.velocity_chord(lm.vx, lm.fx, fz, lm.farea, lm.fmask, False)
        of = fn(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True)
        assert normwise(s["vu"], ov) < 1e-13 and normwise(s["fu"], of) < 1e-13
    ps = O.streamfn_chord(lm.fx, lm.fx, fz, lm.farea, lm.fmask, True)
    assert normwise(s["fpsi"], ps) < 1e-13
    c.close()


@pytest.mark.parametrize("drop,near,lds", [(0, 4.0, True), (0, 2.0, False), (0, 1e30, True), (0, 0.0, True),
                                           (137, 4.0, True), (137, 0.0, False)])
def test_fp64_lagrangian_near_variants(omesh, drop, near, lds):
    """The FP64 Lagrangian-scheme sums take the explicit |y - x|^2 form on tiles
    within `near` tile radii of a target and (1 - x.c) - x.d beyond: every
    threshold (0: the far form on every tile without a self pair; 1e30: the
    explicit form everywhere) and a partly padded last tile (137 of level 5's
    faces removed) agree with the oracle's chord-form sums.  Level 5, 16
    Gaussian vortices plus a constant, so total circulation is nonzero and
    psi's ln 2 term counts."""
    m = omesh(5)
    lm = LeafMesh(m)
    vz, fz = gauss_field(m, seed=6, nvort=16, sigma=0.15)
    fz = fz[m.fmask == 0] + 0.3
    vz = vz + 0.3
    keep = np.ones(len(fz), bool)
    if drop:
        keep[np.random.default_rng(9).choice(len(fz), drop, replace=False)] = False
        assert keep.sum() % 256 != 0
    fx, fzk, fa = lm.fx[keep], fz[keep], lm.farea[keep]
    fm = np.zeros(len(fx), np.uint8)
    c = L.Context(0)
    c.set_state(lm.vx, vz, fx, fzk, fa, fm)
    c.set_source_staging(lds)
    c.set_near_factor(near)
    c.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    s = c.get_state()
    ov = O.velocity_chord(lm.vx, fx, fzk, fa, fm, False)
    of = O.velocity_chord(fx, fx, fzk, fa, fm, True)
    pv = O.streamfn_chord(lm.vx, fx, fzk, fa, fm, False)
    pf = O.streamfn_chord(fx, fx, fzk, fa, fm, True)
    # far pairs of adjacent tiles (near = 0) lose up to ~1e-16 / (h^2 / 2) relative
    tol = 1e-12 if near == 0.0 else 1e-13
    ev = max(normwise(s["vu"], ov), normwise(s["fu"], of))
    ep = max(normwise(s["vpsi"], pv), normwise(s["fpsi"], pf))
    assert ev < tol and ep < tol, (ev, ep)
    c.close()


@pytest.mark.parametrize("level", range(6))
def test_gpu_reproduces_published_poisson_table(omesh, level):
    """SpherePoisson's Y_5^4 problem (src/LpmSpherePoisson.hpp:428-475) through
    the GPU stream-function solve: psi L-inf error at vertices and active faces
    equals the reference's published table (scripts/poissonOutput.py:6-30) to
    the printed six digits."""
    from test_oracle_golden import POISSON_ICOS, sig6
    m = omesh(level)
    fv, _ = O.sphharm54(m.vx)
    ff, _ = O.sphharm54(m.fx)
    c = L.Context(0)
    c.set_state(m.vx, fv, m.fx, ff, m.farea, m.fmask)
    c.solve(L.PREC_FP64, L.SCHEME_REF, want_psi=True)
    s = c.get_state()
    lv = m.fmask == 0
    ev = np.abs(s["vpsi"] - fv / 30).max()
    ef = np.abs(s["fpsi"][lv] - ff[lv] / 30).max()
    assert sig6(ev, POISSON_ICOS[level][1]) and sig6(ef, POISSON_ICOS[level][2]), (ev, ef)
    c.close()


def test_edge_cases(omesh):
    """Faces-only problem (no vertices), a single active source, level 0."""
    m = omesh(0)
    fz = rh_field(m.fx)
    c = L.Context(0)
    c.set_state(np.zeros((0, 3)), np.zeros(0), m.fx, fz, m.farea, m.fmask)
    c.solve(L.PREC_FP64, L.SCHEME_REF, want_psi=True)
    s = c.get_state()
    of = O.velocity(m.fx, m.fx, fz, m.farea, m.fmask, True)
    assert normwise(s["fu"], of) < 1e-13
    c.advance(0.01, TWO_PI, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=2)
    assert np.isfinite(c.get_state()["fx"]).all()
    one = np.ones(m.nf, np.uint8)
    one[5] = 0
    c.set_state(m.vx, rh_field(m.vx), m.fx, fz, m.farea, one)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    s = c.get_state()
    assert np.all(s["fu"][5] == 0)  # the only source excludes itself
    ov = O.velocity_chord(m.vx, m.fx, fz, m.farea, one, False)
    assert normwise(s["vu"], ov) < 1e-6
    c.close()


def test_error_paths():
    c = L.Context(0)
    with pytest.raises(L.LpmError):
        c.solve(L.PREC_FP64, L.SCHEME_REF)  # no state yet
    m = O.Mesh(1)
    with pytest.raises(L.LpmError):  # every face masked: no source
        c.set_state(m.vx, np.zeros(m.nv), m.fx, np.zeros(m.nf), m.farea, np.ones(m.nf, np.uint8))
    c.set_state(m.vx, np.zeros(m.nv), m.fx, np.zeros(m.nf), m.farea, m.fmask)
    with pytest.raises(L.LpmError):  # FP32C needs the renormalised scheme
        c.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP32C)
    with pytest.raises(L.LpmError):
        c.set_launch(5, 0)
    with pytest.raises(L.LpmError):
        c.advance(float("nan"), TWO_PI)
    c.close()
    with pytest.raises(L.LpmError):
        L.Context(99)


def test_cubed_sphere_bve_matches_oracle(tmp_path):
    """BVESphere<CubedSphereSeed> (src/LpmBVESphere.cpp:202): on the level-4
    cubed sphere (1,538 vertices, 1,536 quadrilateral leaves) the FP64
    reference step matches the oracle after 20 steps (velocity 1e-11,
    positions 1e-9), the FP32C Lagrangian evaluation matches the oracle's
    chord-form sums to 1e-6, and a checkpoint names the seed and resumes bit
    for bit."""
    import scipy.io
    m = O.Mesh(4, kind="cubed")
    vz, fz = rh_field(m.vx), rh_field(m.fx)
    lv = m.fmask == 0
    o = O.State(m, vz, fz)
    o.init_velocity(0)
    for _ in range(20):
        o.step(0.01, TWO_PI, 0)
    c = L.Context(0)
    c.set_state(m.vx, vz, m.fx, fz, m.farea, m.fmask)
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    c.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=20)
    s = c.get_state()
    ev = max(normwise(s["vu"], o.vvel), normwise(s["fu"][lv], o.fvel[lv]))
    ex = max(normwise(s["vx"], o.vx), normwise(s["fx"][lv], o.fx[lv]))
    print(f"cubed sphere level 4, 20 steps: velocity {ev:.3e} positions {ex:.3e}")
    assert ev <= 1e-11 and ex <= 1e-9
    path = str(tmp_path / "cubed.nc")
    c.checkpoint(path, time=0.2, omega=TWO_PI, level=4)
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    assert f._attributes["MeshSeed"] == b"CubedSphereSeed" and f.dimensions["nfaceverts"] == 4
    c.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=2)
    b = L.Context(0)
    b.restore(path)
    b.advance(0.01, TWO_PI, L.SCHEME_REF, L.PREC_FP64, nsteps=2)
    sa, sb = c.get_state(), b.get_state()
    for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
        assert np.array_equal(sa[k], sb[k]), k
    c.set_state(m.vx, vz, m.fx, fz, m.farea, m.fmask)
    c.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    s32 = c.get_state()
    ov = O.velocity_chord(m.vx, m.fx, fz, m.farea, m.fmask, False)
    of = O.velocity_chord(m.fx, m.fx, fz, m.farea, m.fmask, True)
    assert normwise(s32["vu"], ov) < 1e-6 and normwise(s32["fu"][lv], of[lv]) < 1e-6
    c.close()
    b.close()


@pytest.mark.parametrize("prec,scheme,tol", [(L.PREC_FP64, L.SCHEME_REF, 1e-13), (L.PREC_FP32C, L.SCHEME_LAGR, 1e-6)])
def test_streamfn_alone_leaves_the_steps_alone(omesh, prec, scheme, tol):
    """lpm_bve_streamfn (the reference driver's post-timestep stream-function
    solve, tests/LpmBVETest.cpp:155-159) between steps: psi matches the oracle
    at the current state, the held velocity is untouched, and the steps after
    it are bit for bit those of a run without it."""
    m = omesh(4)
    lm = LeafMesh(m)
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    a, b = make_ctx(lm, vz, fz), make_ctx(lm, vz, fz)
    for c in (a, b):
        c.advance(0.01, TWO_PI, scheme, prec, nsteps=2)
    before = a.get_state()
    a.streamfn(prec, scheme)
    s = a.get_state()
    for k in ("vx", "vu", "fx", "fu", "fzeta"):
        assert np.array_equal(s[k], before[k]), k
    fn = O.streamfn if scheme == L.SCHEME_REF else O.streamfn_chord
    pv = fn(s["vx"], s["fx"], s["fzeta"], lm.farea, lm.fmask, False)
    pf = fn(s["fx"], s["fx"], s["fzeta"], lm.farea, lm.fmask, True)
    assert normwise(s["vpsi"], pv) < tol and normwise(s["fpsi"], pf) < tol
    for c in (a, b):
        c.advance(0.01, TWO_PI, scheme, prec, nsteps=2)
    sa, sb = a.get_state(), b.get_state()
    for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()
