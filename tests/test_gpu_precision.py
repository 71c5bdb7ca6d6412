"""The precision form at its own configuration: icosahedral level 8 (655,362
vertices, 1,310,720 faces), 32 Gaussian vortices (seed 1, sigma 0.1, area
mean removed) with Omega = 2 pi, 10 Lagrangian RK4 steps of dt = 0.005 in FP32C
(FP32 pair sums, FP64 stage arithmetic, renormalisation and vorticity) and in
FP64 from the same state.  Bounds: velocity 1e-6 max relative (the FP32C
target), stream function 1e-6, positions 1e-8 after the 10 steps."""
import math

import numpy as np
import pytest

import lpmkokkos_amd as L

pytestmark = pytest.mark.gpu

DT, OMEGA, STEPS = 0.005, 2 * math.pi, 10


def test_level8_ten_steps_fp32c_vs_fp64():
    mesh = L.TreeMesh(8)
    lv = mesh.fmask == 0
    fx = np.ascontiguousarray(mesh.fx[lv])
    farea = np.ascontiguousarray(mesh.farea[lv])
    nl = len(fx)
    assert mesh.nverts == 655362 and nl == 1310720
    ic = L.VorticityInitialCondition().with_vortices(1, 32, 1.0, 0.1)
    vz, fz = ic.eval_many(mesh.vx), ic.eval_many(fx)
    mu = L.area_mean(fz, farea, np.zeros(nl, np.uint8))
    vz, fz = vz - mu, fz - mu
    mask = np.zeros(nl, np.uint8)
    c64, c32 = L.Context(0), L.Context(0)
    c64.set_state(mesh.vx, vz, fx, fz, farea, mask)
    c32.set_state(mesh.vx, vz, fx, fz, farea, mask)
    c64.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP64, nsteps=STEPS)
    c32.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=STEPS)
    a, b = c64.get_state(), c32.get_state()
    x64 = np.concatenate([a["vx"], a["fx"]])
    x32 = np.concatenate([b["vx"], b["fx"]])
    assert np.all(np.isfinite(x32))
    dx = np.linalg.norm(x32 - x64, axis=1).max()  # |x| = 1: absolute = relative
    # the particles moved: the comparison is not of two unchanged states
    moved = np.linalg.norm(x64[: mesh.nverts] - mesh.vx, axis=1).max()
    # both precisions at the FP64 end state: velocity and stream function
    c64.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
    a = c64.get_state()
    c32.set_state(a["vx"], a["vzeta"], a["fx"], a["fzeta"], farea, mask)
    c32.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=True)
    b = c32.get_state()
    u64, u32 = np.concatenate([a["vu"], a["fu"]]), np.concatenate([b["vu"], b["fu"]])
    p64, p32 = np.concatenate([a["vpsi"], a["fpsi"]]), np.concatenate([b["vpsi"], b["fpsi"]])
    du = np.linalg.norm(u32 - u64, axis=1).max() / np.linalg.norm(u64, axis=1).max()
    dp = np.abs(p32 - p64).max() / np.abs(p64).max()
    print(f"level 8, {STEPS} steps: positions {dx:.3e} (moved {moved:.3e}), velocity {du:.3e}, psi {dp:.3e}")
    assert moved > 1e-3
    assert dx < 1e-8 and du < 1e-6 and dp < 1e-6
    c64.close()
    c32.close()
