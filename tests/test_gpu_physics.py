"""Physical checks on the GPU path: the Rossby-Haurwitz wave's pattern speed,
and the C++ driver (the reference's BVE test driver on this library)."""
import math
import os
import subprocess

import numpy as np
import pytest

import lpmkokkos_amd as L
from helpers import LeafMesh, rh_field

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rh_phase(x, zeta, n=4, lat0=0.6, band=0.15):
    lat = np.arctan2(x[:, 2], np.hypot(x[:, 0], x[:, 1]))
    lon = np.arctan2(x[:, 1], x[:, 0])
    sel = np.abs(lat - lat0) < band
    c = np.sum(zeta[sel] * np.cos(n * lon[sel]))
    s = np.sum(zeta[sel] * np.sin(n * lon[sel]))
    return math.atan2(s, c) / n


@pytest.mark.parametrize("prec", [L.PREC_FP32C, L.PREC_FP64])
def test_rossby_haurwitz_pattern_speed(omesh, prec):
    """RH n=4, w=K=0.1, Omega=2 pi: the wave pattern drifts at
    nu = (n(n+3)w - 2 Omega)/((n+1)(n+2)) = -0.3255 rad per unit time."""
    lm = LeafMesh(omesh(5))
    vz, fz = rh_field(lm.vx), rh_field(lm.fx)
    c = L.Context(0)
    c.set_state(lm.vx, vz, lm.fx, fz, lm.farea, lm.fmask)
    p0 = rh_phase(lm.fx, fz)
    dt, n = 0.01, 100
    c.advance(dt, 2 * math.pi, L.SCHEME_LAGR, prec, nsteps=n)
    s = c.get_state()
    d = rh_phase(s["fx"], s["fzeta"]) - p0
    d = (d + math.pi / 4) % (math.pi / 2) - math.pi / 4
    nu = L.RossbyHaurwitzWave(4, 0.1, 0.1).phase_speed(2 * math.pi)
    print(f"RH pattern speed {d / (dt * n):.6f} vs analytic {nu:.6f}")
    assert abs(d / (dt * n) - nu) < 0.02 * abs(nu)
    c.close()


@pytest.mark.parametrize("args", [["-d", "3", "-seed", "icos"],
                                  ["-d", "5", "-seed", "icos", "-ic", "rh", "-scheme", "lagr", "-prec", "fp32c"],
                                  ["-d", "4", "-seed", "icos", "-scheme", "lagr", "-f", "50"],
                                  ["-d", "4"],
                                  ["-d", "6", "-ic", "rh", "-scheme", "lagr", "-prec", "fp32c"],
                                  ["-d", "5", "-scheme", "lagr", "-f", "50"]])
def test_cpp_driver(args):
    """The reference's BVE driver on this library: solid-body rotation errors
    and the Rossby-Haurwitz pattern speed within the driver's bounds, on the
    cubed sphere (the driver's default seed, tests/LpmBVETest.cpp:42) and on
    the icosahedral mesh."""
    exe = os.path.join(REPO, "lpmkokkos_amd", "bve_driver")
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tests pass" in r.stdout
