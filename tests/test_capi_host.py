"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/lpm_bve.h declares, and its host pieces (mesh build, initial
conditions) match the oracle bit for bit."""
import ctypes as C
import re

import numpy as np
import pytest

import lpmkokkos_amd as L
from lpmkokkos_amd import _lib


def declared_symbols():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(lpm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 25
    lib = _lib.lib()
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_error_reporting():
    lib = _lib.lib()
    nv = C.c_int64()
    rc = lib.lpm_icos_alloc(-1, C.byref(nv), None, None)
    assert rc == 1 and b"level" in lib.lpm_last_error()
    assert lib.lpm_icos_alloc(3, C.byref(nv), None, None) == 0 and nv.value == 642
    assert lib.lpm_last_error() == b""
    with pytest.raises(L.LpmError):
        L.TreeMesh(13)
    rc = lib.lpm_icos_tree_init(2, b"/nonexistent/seed.dat", None, None, None, None, None, None, None, None, None)
    assert rc == 6


def test_seed_fixture_matches_builtin(oracle_mod):
    from oracle.oracle import SEED_FIXTURE
    for level in (0, 3):
        a = L.TreeMesh(level)
        b = L.TreeMesh(level, L.MeshSeed(path=SEED_FIXTURE))
        for k in ("vx", "fx", "farea", "fmask", "fverts", "fparent"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4, 5, 6])
def test_mesh_bitwise_equals_oracle(omesh, level):
    m = L.TreeMesh(level)
    o = omesh(level)
    assert (m.nverts, m.nedges, m.nfaces, m.nleaves) == (o.nv, o.ne, o.nf, o.nleaves)
    for a, b in ((m.vx, o.vx), (m.vlag, o.vlag), (m.fx, o.fx), (m.flag, o.flag), (m.farea, o.farea),
                 (m.fmask, o.fmask), (m.fverts, o.fverts), (m.fparent, o.fparent)):
        assert np.array_equal(a, b)


def test_ic_bitwise_equals_oracle(oracle_mod, omesh):
    O = oracle_mod
    m = omesh(4)
    c, a = L.gauss_params(7, 32, 1.0)
    co, ao = O.gauss_params(7, 32, 1.0)
    assert np.array_equal(c, co) and np.array_equal(a, ao)
    assert np.allclose(np.linalg.norm(c, axis=1), 1.0, atol=1e-15)
    ic = L.VorticityInitialCondition(sbr=0.5, rh_n=4, rh_w=0.1, rh_K=0.1, centers=c, amps=a, sigma=0.1)
    for x in (m.vx, m.fx):
        z = ic.eval_many(x)
        zo = O.ic_eval(x, sbr=0.5, rh_n=4, rh_w=0.1, rh_K=0.1, centers=co, amps=ao, sigma=0.1)
        assert np.array_equal(z, zo)
    sbr = L.SolidBodyRotation()
    assert sbr.eval(0.0, 0.6, 0.8) == 2 * (2 * L.bve.PI) * 0.8


def test_area_mean(omesh):
    m = omesh(3)
    z = np.arange(m.nf, dtype=np.float64)
    mu = L.area_mean(z, m.farea, m.fmask)
    lv = m.fmask == 0
    assert abs(mu - (z[lv] * m.farea[lv]).sum() / m.farea[lv].sum()) < 1e-9


def test_rh_phase_speed():
    ic = L.RossbyHaurwitzWave(4, 0.1, 0.1)
    # nu = (n(n+3) w - 2 Omega) / ((n+1)(n+2)) = (2.8 - 4 pi) / 30
    assert abs(ic.phase_speed(2 * np.pi) - (2.8 - 4 * np.pi) / 30) < 1e-15


def test_config3_members_build():
    from lpmkokkos_amd import ensemble as E
    sp = E.ensemble_config(seed=3, n=64)
    assert len(sp) == 64 and {s.level for s in sp} <= {3, 4, 5, 6, 7}
    assert all(0.95 <= s.k_factor <= 1.05 for s in sp)
    assert len({s.vortex_seed for s in sp}) == 64
    u, n = E.rng_draws(9, 1000, 1000)
    assert 0 <= u.min() and u.max() < 1 and abs(n.mean()) < 0.15
    assert E.state_bytes(655362, 1310720) > 29 * 8 * (655362 + 1310720)


@pytest.mark.parametrize("level,splits", [(4, 5), (5, 20), (6, 64), (7, 72), (8, 69), (9, 18)])
def test_state_bytes_count_the_default_split_partials(level, splits):
    """The memory a context plans (and serving admission charges) holds the
    default launch's split partials: 4 FP64 sums per target and source split,
    with short workgroups: 64 and 72 splits at levels 6 and 7, 69 at level 8, 18 at
    level 9, where the 4.6 GB bound on the partials sets the count, and at
    least 4 tiles per split below (5 and 20 at levels 4 and 5)."""
    from lpmkokkos_amd import ensemble as E
    nv, na = 2 + 10 * 4 ** level, 20 * 4 ** level
    nt = nv + na
    ntiles = -(-na // 256)
    ns_pad = 256 * ntiles
    state = 8 * nt * 30 + 4 * (nt + na)
    records = ns_pad * (64 + 32 + 8) + (4 * ntiles * 32 + 256) + (4 * ntiles + 256)
    assert E.state_bytes(nv, na) == state + records + splits * 4 * nt * 8 + 8 * 256


def test_state_bytes_rejects_sizes_set_state_rejects():
    from lpmkokkos_amd import ensemble as E
    nv, na = 2 + 10 * 4 ** 14, 20 * 4 ** 14  # beyond lpm_bve_set_state's range
    with pytest.raises(L.LpmError):
        E.state_bytes(nv, na)


def test_request_generator():
    from lpmkokkos_amd import serve as S
    reqs = S.make_requests(2000, rate=0.5, seed=4)
    lv = np.array([r.spec.level for r in reqs])
    frac = {l: float(np.mean(lv == l)) for l in (4, 5, 6, 7, 8)}
    for l, p in S.LEVEL_PROBS.items():
        assert abs(frac[l] - p) < 0.04, (l, frac[l])
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and abs(np.mean(np.diff(arr)) - 2.0) < 0.2
    assert {r.spec.rh_n for r in reqs} == {2, 3, 4, 5, 6}
    assert min(r.spec.nvort for r in reqs) == 4 and max(r.spec.nvort for r in reqs) == 16


def test_request_sizes_both_seeds():
    """The serving loop's vertex/leaf counts and pairs per step agree with the
    trees the library builds, for both seeds."""
    from lpmkokkos_amd import serve as S
    for seed, ms in ((L.SEED_ICOS_TRI, None), (L.SEED_CUBED_SPHERE, L.MeshSeed(L.CubedSphereSeed))):
        for level in range(4):
            m = L.TreeMesh(level, ms)
            nl = int(np.sum(m.fmask == 0))
            assert S.tree_sizes(level, seed) == (m.nverts, nl), (seed, level)
            assert S.pairs_per_step(level, seed) == 4.0 * ((m.nverts + nl) * nl - nl)


def test_poisson_y54_matches_oracle(oracle_mod):
    """Host Y_5^4 right-hand side and RH54 velocity (src/LpmSpherePoisson.hpp:60-80,
    src/LpmRossbyWaves.hpp:18-53) against the oracle's restatement: velocity bit
    for bit; f to 2 ulp (the oracle's compiler may merge cos(4 lon) with the
    velocity's sin(4 lon) into one sincos call, which rounds differently)."""
    m = L.TreeMesh(3)
    for x in (m.vx, m.fx):
        f, u = L.poisson_y54(x)
        of, ou = oracle_mod.sphharm54(x)
        assert np.all(np.abs(f - of) <= 2 * np.spacing(np.abs(of)))
        assert np.array_equal(u, ou)


def test_err_norms_match_oracle(oracle_mod, omesh):
    """The product's ErrNorms (host) equals the oracle's restatement of
    src/LpmErrorNorms.cpp:15-70 bit for bit, on the solid-body velocity of a
    level-3 mesh with area weights, and the closed forms hold."""
    O = oracle_mod
    m = omesh(3)
    lv = m.fmask == 0
    x = m.fx[lv]
    Om = 2 * np.pi
    exact = np.stack([-Om * x[:, 1], Om * x[:, 0], 0 * x[:, 0]], axis=1)
    rng = np.random.default_rng(5)
    appx = exact + 1e-3 * rng.normal(size=exact.shape)
    w = m.farea[lv]
    n = L.ErrNorms.vector(appx, exact, w)
    assert [n.l1, n.l2, n.linf] == list(O.err_norms_vec(appx, exact, w))
    assert np.array_equal(n.err, appx - exact)
    s_ex = exact[:, 0]
    s_app = appx[:, 0]
    ns = L.ErrNorms.scalar(s_app - s_ex, s_ex, w)
    assert [ns.l1, ns.l2, ns.linf] == list(O.err_norms_scalar(s_app, s_ex, w))
    c = L.ErrNorms.vector(1.25 * exact, exact, w)
    assert abs(c.l1 - 0.25) < 1e-15 and abs(c.l2 - 0.25) < 1e-15 and abs(c.linf - 0.25) < 1e-15
    xs = rng.uniform(1.0, 2.0, 50)
    neg = L.ErrNorms.scalar(-0.5 * xs, xs, np.ones(50))
    assert abs(neg.linf + 0.5 * xs.min() / xs.max()) < 1e-15  # signed maxima, as the reference
    assert n.infoString("velocity").startswith("velocity ErrNorms: l1 = ")
    with pytest.raises(ValueError):
        L.ErrNorms.scalar(np.zeros(3), np.zeros(4), np.ones(3))


def test_velocity_tangent():
    """SphereVelocityTangentTestFunctor: u.x, zero for Omega z x x."""
    m = L.TreeMesh(2)
    u = np.stack([-m.vx[:, 1], m.vx[:, 0], 0 * m.vx[:, 0]], axis=1)
    d = L.velocity_tangent(m.vx, u)
    assert np.abs(d).max() < 1e-15
    d2 = L.velocity_tangent(m.vx, m.vx)
    assert np.abs(d2 - 1).max() < 1e-15


def test_server_rejects_negative_large_lanes():
    """Checked before any device work, so it runs without a GPU."""
    from lpmkokkos_amd import serve as S
    with pytest.raises(ValueError):
        S.Server(large_lanes=-1)
