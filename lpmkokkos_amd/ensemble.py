"""Ensembles of independent BVE problems, launch-shape autotuning, and the
ensemble configuration builder (batching form).

Each of the four evaluations of an ensemble step is one persistent pair
kernel over the flattened (member, target block, source split) work list
(see include/lpm_bve.h); members keep the decomposition they have alone, so
each member's result is bit-identical alone and in the ensemble.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PREC_FP32C, SCHEME_LAGR, SEED_CUBED_SPHERE, SEED_ICOS_TRI, call
from .bve import Context, CubedSphereSeed, MeshSeed, TreeMesh, VorticityInitialCondition, area_mean


def rng_draws(seed: int, nuniform: int = 0, nnormal: int = 0):
    u = np.zeros(max(nuniform, 1))
    n = np.zeros(max(nnormal, 1))
    call("lpm_rng_draws", seed, nuniform, _lib.ptr(u), nnormal, _lib.ptr(n))
    return u[:nuniform], n[:nnormal]


class Ensemble:
    """lpm_bve_ensemble: borrows member contexts (keep them alive)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("lpm_ens_create", device, C.byref(h))
        self.h = h
        self.members: list[Context] = []

    def close(self):
        if self.h:
            _lib.lib().lpm_ens_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, ctx: Context):
        call("lpm_ens_add", self.h, ctx.h)
        self.members.append(ctx)

    def remove(self, ctx: Context):
        call("lpm_ens_remove", self.h, ctx.h)
        self.members.remove(ctx)

    def __len__(self):
        n = C.c_int()
        call("lpm_ens_size", self.h, C.byref(n))
        return n.value

    def set_blocks_per_cu(self, n: int):
        """Workgroups per CU of the persistent kernel (0: the autotuned rule for
        the ensemble's total N, else the occupancy limit)."""
        call("lpm_ens_set_blocks_per_cu", self.h, n)

    def autotune(self, prec: int = PREC_FP32C, scheme: int = SCHEME_LAGR, reps: int = 2, max_blocks: int = 16):
        """Time one evaluation of every member per persistent grid (1 .. the
        occupancy limit of workgroups per CU); the fastest becomes the rule for
        ensembles of this total N (process-wide).  Returns (workgroups per CU,
        ms, {workgroups per CU: ms} of the candidates tried)."""
        best, ms = C.c_int(), C.c_double()
        times = np.zeros(max_blocks)
        call("lpm_ens_autotune", self.h, prec, scheme, reps, C.byref(best), C.byref(ms), _lib.ptr(times),
             max_blocks)
        return best.value, ms.value, {k + 1: float(t) for k, t in enumerate(times) if t >= 0}

    def autotune_shapes(self, prec: int = PREC_FP32C, scheme: int = SCHEME_LAGR, reps: int = 1,
                        work_fraction: float = 0.9) -> int:
        """Launch shapes tuned on whole-ensemble evaluations for the member
        sizes N carrying the most pair work (until work_fraction of it is
        covered); each becomes the process-wide rule for its N, which contexts
        run alone take too.  Returns the number of sizes tuned."""
        n = C.c_int()
        call("lpm_ens_autotune_shapes", self.h, prec, scheme, reps, float(work_fraction), C.byref(n))
        return n.value

    def advance(self, dt: float, omega: float, scheme: int = SCHEME_LAGR, prec: int = PREC_FP32C, nsteps: int = 1,
                use_graph: bool = True):
        call("lpm_ens_advance", self.h, dt, omega, scheme, prec, nsteps, int(use_graph))

    def synchronize(self):
        call("lpm_ens_synchronize", self.h)

    def stream(self) -> int:
        """The ensemble's hipStream_t (for members created on it)."""
        st = C.c_void_p()
        call("lpm_ens_stream", self.h, C.byref(st))
        return st.value

    def done(self) -> bool:
        """True when the steps queued by the ensemble have finished (non-blocking)."""
        d = C.c_int()
        call("lpm_ens_query", self.h, C.byref(d))
        return bool(d.value)


def autotune(ctx: Context, prec: int = PREC_FP32C, scheme: int = SCHEME_LAGR, reps: int = 2):
    """Time the candidate launch shapes on ctx's problem; the fastest becomes
    the rule for its N (process-wide).  Returns (targets_per_lane, splits, ms)."""
    R, S, ms = C.c_int(), C.c_int(), C.c_double()
    call("lpm_bve_autotune", ctx.h, prec, scheme, reps, C.byref(R), C.byref(S), C.byref(ms))
    return R.value, S.value, ms.value


def shape_rule(ntargets: int):
    R, S = C.c_int(), C.c_int()
    if _lib.lib().lpm_shape_rule_get(ntargets, C.byref(R), C.byref(S)) != 0:
        return None
    return R.value, S.value


def state_bytes(nv: int, nactive: int) -> int:
    b = C.c_int64()
    call("lpm_bve_state_bytes", nv, nactive, C.byref(b))
    return b.value


@dataclass
class MemberSpec:
    """One ensemble member or forecast request: mesh (level of the icosahedral
    tree, or of the cubed-sphere tree with seed = SEED_CUBED_SPHERE) and
    initial field (RH wave of wavenumber rh_n with K = 0.1 * k_factor, w = 0.1,
    plus nvort Gaussian vortices of amplitude N(0, amp) and width sigma, mean
    removed)."""
    level: int
    k_factor: float = 1.0
    rh_n: int = 4
    nvort: int = 8
    amp: float = 0.05
    sigma: float = 0.15
    vortex_seed: int = 0
    seed: int = SEED_ICOS_TRI


_MESHES: dict[tuple[int, int], TreeMesh] = {}


def leaf_mesh(level: int, seed: int = SEED_ICOS_TRI):
    """Vertices and active faces of a level-`level` tree of seed `seed`
    (cached per seed and level)."""
    if (seed, level) not in _MESHES:
        _MESHES[(seed, level)] = TreeMesh(level, MeshSeed(CubedSphereSeed) if seed == SEED_CUBED_SPHERE else None)
    m = _MESHES[(seed, level)]
    lv = m.fmask == 0
    return m, np.ascontiguousarray(m.fx[lv]), np.ascontiguousarray(m.farea[lv])


def member_ic(spec: MemberSpec) -> VorticityInitialCondition:
    """RH wave (w = 0.1, K = 0.1 k_factor) plus the member's seeded vortices."""
    ic = VorticityInitialCondition(rh_n=spec.rh_n, rh_w=0.1, rh_K=0.1 * spec.k_factor)
    if spec.nvort:
        ic.with_vortices(spec.vortex_seed, spec.nvort, spec.amp, spec.sigma)
    return ic


def member_state(spec: MemberSpec):
    """Host-built state (leaf faces only), mean removed."""
    m, fx, farea = leaf_mesh(spec.level, spec.seed)
    ic = member_ic(spec)
    vz = ic.eval_many(m.vx)
    fz = ic.eval_many(fx)
    mu = area_mean(fz, farea, np.zeros(len(fz), np.uint8))
    return m.vx, vz - mu, fx, fz - mu, farea


def make_member(spec: MemberSpec, device: int = 0, on_device: bool = True, stream=None) -> Context:
    """A context holding the member's initial state, mean removed: built on the
    device from the cached level mesh (default), or on the host and uploaded.
    stream: queue the context's work there (e.g. its ensemble's stream)."""
    c = Context(device, stream=stream)
    if on_device:
        c.set_state_level(spec.level, member_ic(spec), subtract_mean=True, seed_kind=spec.seed)
        return c
    vx, vz, fx, fz, farea = member_state(spec)
    c.set_state(vx, vz, fx, fz, farea, np.zeros(len(fz), np.uint8))
    return c


def ensemble_config(seed: int = 3, n: int = 64, levels=(3, 4, 5, 6, 7)):
    """Batching configuration: n members, level uniform on `levels`, K scaled
    by U(0.95, 1.05), 8 Gaussian vortices of amplitude N(0, 0.05), sigma 0.15,
    all drawn from `seed`."""
    u, _ = rng_draws(seed, 2 * n)
    specs = []
    for m in range(n):
        lev = levels[min(len(levels) - 1, int(u[2 * m] * len(levels)))]
        specs.append(MemberSpec(level=lev, k_factor=0.95 + 0.1 * u[2 * m + 1], vortex_seed=1000 * seed + m))
    return specs
