import numpy as np

from helpers import LeafMesh, rh_field
from oracle import oracle as O


def test_leaf_restriction_is_exact(omesh):
    """The oracle over active faces only reproduces the masked tree loop bit for
    bit at vertices and active faces, over RK4 steps too."""
    m = omesh(3)
    lm = LeafMesh(m)
    vz, fz = rh_field(m.vx), rh_field(m.fx)
    a = O.State(lm, vz, fz[m.fmask == 0])
    b = O.State(m, vz, fz)
    a.init_velocity(0)
    b.init_velocity(0)
    for _ in range(2):
        a.step(0.01, 2 * np.pi, 0)
        b.step(0.01, 2 * np.pi, 0)
    lv = m.fmask == 0
    assert np.array_equal(a.vx, b.vx) and np.array_equal(a.fx, b.fx[lv])
    assert np.array_equal(a.fzeta, b.fzeta[lv]) and np.array_equal(a.fvel, b.fvel[lv])
