#!/usr/bin/env python3
"""Benchmark: spherical Biot-Savart pair interactions per second on one MI355X.

Workload (the metric's configuration, 1.31M sources): icosahedral level 8 --
655,362 vertex and 1,310,720 face targets, 1,310,720 face sources -- with the
sum of 32 seeded Gaussian vortices (sigma 0.1, amplitudes N(0,1), mean
removed) over an Omega = 2 pi background, advanced by the renormalised
Lagrangian RK4 step (dt = 0.005).  A step is one advance_timestep: four
direct-sum evaluations over all targets (stages 2-4 and the end-of-step
solve, which is the next step's stage 1).

value  = FP32C pair interactions per second over the K timed steps, whole
         job, state resident in HBM, steps replayed from a captured hipGraph.
Also reported, in the same line:
  * the FP64 path's rate over 10 steps from the same initial state, the
    FP32C/FP64 speed-up beside its instruction-issue bound, FP32C error
    against FP64 (positions after those 10 steps; velocity from the step
    kernel and stream function at the final FP64 state), the FP64 reference
    scheme's rate, and the pair kernel's roofline position (HIP events around
    the kernel on its own stream);
  * the reference-side copy-in/copy-out binding of advance_timestep on host
    arrays (one FP64 reference-scheme step per call): its PCIe-inclusive rate
    and the time its copies add to the resident step;
  * one timed FP64 reference RK4 step at level 7 (the same field);
  * the batching configuration in full (64 members, levels 3-7, 10 FP32C
    steps): ensemble rate, the members' sequential rate, and whether every
    member equals its run alone bit for bit;
  * a 40-request sample of the serving configuration (levels 4-8 with the
    configuration's probabilities, 20 steps each): p50/p99 overall and by
    level, aggregate pair interactions/s;
  * the CPU baseline: the oracle's reference-structured loop on the host's
    physical cores (C0 and C1 in full, C2 on a 1/64 target subset).

Pairs per evaluation = (nv + nl) * nl - nl (target x active source, self
excluded).  Algorithmic flops per pair = 15 (d = y - x: 3, |d|^2: 5,
q/|d|^2: 1, accumulate s*d: 6).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# one hardware queue per lane stream of the serving sample (the runtime reads
# this at the first HIP call; its default is 4)
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FLOPS_PER_PAIR = 15
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
# HBM bytes per level-8 FP32C pair launch at the default launch shape: FETCH_SIZE + WRITE_SIZE
# from two rocprofv3 --pmc runs, each scaled by the same counter's reading for a device copy of
# known size
TRAFFIC_BYTES_PER_LAUNCH = 4.1810e9 + 4.3496e9
# FP32C/FP64 speed-up per step allowed by the pair loops' instruction issue
# (22.2 vs 45.0 SIMD cycles per 64 pairs, measured issue costs)
FP32C_VS_FP64_ISSUE_BOUND = 45.0 / 22.2
DT = 0.005
OMEGA = 2 * np.pi


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class ClockSampler:
    """GFX clock of every XCD and socket power, read with amdsmi every
    `interval` seconds in a background thread while the timed steps run.  The
    published peaks assume 2,400 MHz; boxes run this load at lower clocks.
    Any failure (no amdsmi, no access) leaves the figures out."""

    def __init__(self, interval: float = 0.25):
        import threading
        self.interval = interval
        self.samples = []  # (mean XCD clock, min, max, socket power)
        self.error = None
        self._stop = threading.Event()
        self._thread = threading.Thread(target=self._run, daemon=True)

    def _run(self):
        try:
            import amdsmi
            amdsmi.amdsmi_init()
            h = amdsmi.amdsmi_get_processor_handles()[0]
            while not self._stop.is_set():
                m = amdsmi.amdsmi_get_gpu_metrics_info(h)
                clks = [c for c in m.get("current_gfxclks", []) if isinstance(c, (int, float)) and c > 0]
                pw = m.get("current_socket_power")
                if clks:
                    self.samples.append((sum(clks) / len(clks), min(clks), max(clks),
                                         pw if isinstance(pw, (int, float)) else None))
                self._stop.wait(self.interval)
            amdsmi.amdsmi_shut_down()
        except Exception as e:  # reported, never fatal
            self.error = repr(e)

    def start(self):
        self._thread.start()
        return self

    def stop(self):
        self._stop.set()
        self._thread.join(timeout=2.0)
        if not self.samples:
            return {"error": self.error or "no samples"}
        n = len(self.samples)
        pws = [x[3] for x in self.samples if x[3] is not None]
        return {
            "gfx_mhz_mean": sum(x[0] for x in self.samples) / n,
            "gfx_mhz_min_xcd": min(x[1] for x in self.samples),
            "gfx_mhz_max_xcd": max(x[2] for x in self.samples),
            "socket_power_w_mean": sum(pws) / len(pws) if pws else None,
            "samples": n,
            "source": f"amdsmi current_gfxclks (8 XCDs) and current_socket_power every {self.interval} s "
                      "during the timed FP32C steps",
        }


def drop_in_binding(ctx, farea, nl: int, pairs_step: int, resident_s: float, calls: int = 2):
    """The reference-side binding of advance_timestep on host arrays: copy the
    state and entry velocity in (lpm_bve_set_state + lpm_bve_set_velocity, in
    place after the first call), one FP64 reference-scheme step, copy the state
    out (lpm_bve_get_state) -- the PCIe-inclusive rate of the drop-in path.
    One untimed call, then `calls` timed ones."""
    from lpmkokkos_amd import _lib
    import lpmkokkos_amd as L
    s = ctx.get_state()
    mask = np.zeros(nl, np.uint8)
    ts = []
    for i in range(calls + 1):
        t = time.perf_counter()
        ctx.set_state(s["vx"], s["vzeta"], s["fx"], s["fzeta"], farea, mask)
        ctx.set_velocity(s["vu"], s["fu"])
        ctx.advance(DT, OMEGA, L.SCHEME_REF, L.PREC_FP64, nsteps=1)
        _lib.call("lpm_bve_get_state", ctx.h, _lib.ptr(s["vx"]), _lib.ptr(s["vzeta"]), _lib.ptr(s["vu"]), None,
                  _lib.ptr(s["fx"]), _lib.ptr(s["fzeta"]), _lib.ptr(s["fu"]), None)
        if i:
            ts.append(time.perf_counter() - t)
    t = min(ts)
    return {
        "workload": "one FP64 reference-scheme step per call on host arrays: set_state + set_velocity, "
                    "advance, get_state (state in place on the device after the first call)",
        "ms_per_step": 1e3 * t,
        "resident_ms_per_step": 1e3 * resident_s,
        "copy_ms_per_step": 1e3 * (t - resident_s),
        "pairs_per_s": pairs_step / t,
        "calls_timed": calls,
    }


def build_problem(level: int, seed: int = 1):
    import lpmkokkos_amd as L
    mesh = L.TreeMesh(level)
    lv = mesh.fmask == 0
    fx = np.ascontiguousarray(mesh.fx[lv])
    farea = np.ascontiguousarray(mesh.farea[lv])
    ic = L.VorticityInitialCondition().with_vortices(seed, 32, 1.0, 0.1)
    vz = ic.eval_many(mesh.vx)
    fz = ic.eval_many(fx)
    mu = L.area_mean(fz, farea, np.zeros(len(fz), np.uint8))
    return mesh, fx, farea, vz - mu, fz - mu


def cpu_baseline(timeout_s: float = 240.0):
    """The oracle's reference-structured loop (oracle/cpu_baseline.py) on the
    host's physical cores, in a child process so that its OpenMP thread
    binding leaves this process's threads alone: C0 and C1 in full, C2 on a
    1/64 strided target subset."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py")], env=env,
                       capture_output=True, text=True, timeout=timeout_s)
    sys.stderr.write(r.stderr)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed ({r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def serving_form(nreq: int, max_level: int, load: float = 0.8, steps: int = 20, seed: int = 4):
    """Continuous-batching sample of the serving configuration: Poisson
    arrivals at `load` x measured capacity, levels from {4..max_level} with the
    configuration's probabilities (renormalised), RH n in 2..6 + 4..16
    Gaussian vortices, 20 FP32C steps each, 200 GB admission budget."""
    from lpmkokkos_amd import serve as S
    probs = {l: p for l, p in S.LEVEL_PROBS.items() if l <= max_level}
    tot = sum(probs.values())
    probs = {l: p / tot for l, p in probs.items()}
    svc = S.measure_service_times(sorted(probs), steps=steps)
    mean_service = sum(probs[l] * svc[l] for l in probs)
    rate = load / mean_service
    reqs = S.make_requests(nreq, rate, seed=seed, level_probs=probs, steps=steps)
    srv = S.Server(budget_bytes=200e9, log=log)
    res = srv.run(reqs)
    srv.close()
    levels = res["levels"]
    res.update({"arrival_rate_per_s": rate, "load": load, "level_probs": probs,
                "service_s_by_level": {int(k): v for k, v in svc.items()}, "steps_per_request": steps,
                "sample": f"{nreq} requests of the 2000-request serving configuration (seed {seed}; levels "
                          + ", ".join(f"{l}: {n}" for l, n in sorted(levels.items()))
                          + f"), Poisson arrivals at {load:.0%} of the measured capacity; p99 of {nreq} "
                          "completions is close to their maximum"})
    return res


def ensemble_run(nmembers: int = 64, nsteps: int = 10) -> dict:
    """Batching configuration: 64 members, levels uniform on 3..7 (seed 3), RH
    wave (K x U(0.95, 1.05)) + 8 Gaussian vortices, 10 FP32C steps, one
    persistent pair kernel per evaluation, step graph replayed, launch shapes
    from the autotuner (per level alone, then re-tuned on the whole ensemble
    for the levels that carry 90 % of the pairs) and the persistent grid
    (workgroups per CU) from the grid autotuner.  Every member is then advanced alone by the same 10
    steps (one at a time, as a caller without batching would run them: the
    sequential rate) and compared with its ensemble result bit for bit."""
    import lpmkokkos_amd as L
    from lpmkokkos_amd import ensemble as E
    specs = E.ensemble_config(seed=3, n=nmembers)
    t0 = time.perf_counter()
    members = [E.make_member(s) for s in specs]
    tuned = {}
    for s, c in zip(specs, members):
        if s.level not in tuned:
            tuned[s.level] = E.autotune(c, reps=1)
    log(f"[bench] ensemble: {len(members)} members built and shapes autotuned in {time.perf_counter() - t0:.1f} s: "
        f"{tuned}")
    ens = E.Ensemble(0)
    for c in members:
        ens.add(c)
    # shapes of the sizes carrying 90 % of the pairs re-tuned on whole-ensemble evaluations
    # (the rule for N, which the members alone below take too)
    ntuned = ens.autotune_shapes(L.PREC_FP32C, L.SCHEME_LAGR, reps=1, work_fraction=0.9)
    ens_shapes = {s.level: E.shape_rule(c.sizes()[3]) for s, c in zip(specs, members)}
    log(f"[bench] ensemble: shapes of {ntuned} size(s) re-tuned on the ensemble: {ens_shapes}")
    # waves per CU of the persistent kernel: the grid autotuner's rule for this total N
    grid = ens.autotune(L.PREC_FP32C, L.SCHEME_LAGR, reps=1)
    log(f"[bench] ensemble: persistent grid autotuned: {grid[0]} workgroups per CU "
        f"(ms per evaluation by workgroups per CU: {', '.join(f'{k}: {v:.1f}' for k, v in grid[2].items())})")
    ens.advance(DT, OMEGA, nsteps=1)  # warm: entry velocities, plan, capture
    ens.synchronize()
    t1 = time.perf_counter()
    ens.advance(DT, OMEGA, nsteps=nsteps)
    ens.synchronize()
    dt_ens = time.perf_counter() - t1
    pairs_step = sum(4.0 * ((2 + 10 * 4 ** s.level + 20 * 4 ** s.level) * 20 * 4 ** s.level - 20 * 4 ** s.level)
                     for s in specs)
    log(f"[bench] ensemble: {1e3 * dt_ens / nsteps:.1f} ms/step, {pairs_step * nsteps / dt_ens:.4e} pairs/s")
    dt_seq = 0.0
    same = 0
    for s, b in zip(specs, members):
        a = E.make_member(s)
        a.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=1)  # the warm step, as the ensemble's
        a.synchronize()
        t = time.perf_counter()
        a.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=nsteps)
        a.synchronize()
        dt_seq += time.perf_counter() - t
        sa, sb = a.get_state(), b.get_state()
        same += all(np.array_equal(sa[k], sb[k]) for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"))
        a.close()
    log(f"[bench] ensemble: members alone {pairs_step * nsteps / dt_seq:.4e} pairs/s; {same}/{len(specs)} "
        "bitwise equal")
    ens.close()
    for c in members:
        c.close()
    return {
        "workload": f"{len(specs)} members, levels uniform on 3..7 (seed 3), RH n=4 K x U(0.95,1.05) "
                    "+ 8 Gaussian vortices N(0,0.05) sigma 0.15, Lagrangian RK4 dt=0.005, FP32C",
        "levels": {int(l): sum(1 for s in specs if s.level == l) for l in sorted({s.level for s in specs})},
        "steps": nsteps,
        "pairs_per_step": pairs_step,
        "ms_per_step": 1e3 * dt_ens / nsteps,
        "pairs_per_s": pairs_step * nsteps / dt_ens,
        "members_sequential_pairs_per_s": pairs_step * nsteps / dt_seq,
        "speedup_vs_sequential": dt_seq / dt_ens,
        "members_bitwise_equal_alone": f"{same}/{len(specs)}",
        "all_members_bitwise_equal_alone": same == len(specs),
        "autotuned_shapes": {int(k): list(v) for k, v in tuned.items()},
        "ensemble_tuned_shapes": {int(k): list(v) for k, v in ens_shapes.items() if v},
        "autotuned_workgroups_per_cu": grid[0],
        "eval_ms_by_workgroups_per_cu": grid[2],
    }


def ensemble_form(args) -> int:
    r = ensemble_run(args.ensemble_members)
    out = {
        "metric": "ensemble pair-interactions/s (batching form)",
        "value": r["pairs_per_s"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": r["workload"], "levels": r["levels"], "pairs_per_step": r["pairs_per_step"]},
        "ensemble": r,
    }
    print(json.dumps(out), flush=True)
    return 0


def c1_gpu_step(level: int = 7) -> dict:
    """Plain FP64 BVE at level 7 (163,842 vertices, 327,680 faces), the
    32-Gaussian-vortex field of seed 1: one timed RK4 step of the reference
    scheme (literal 1 - x.y, dt = 0.01) after one untimed step."""
    import lpmkokkos_amd as L
    mesh, fx, farea, vz, fz = build_problem(level)
    nv, nl = mesh.nverts, len(fx)
    c = L.Context(0)
    c.set_state(mesh.vx, vz, fx, fz, farea, np.zeros(nl, np.uint8))
    c.solve(L.PREC_FP64, L.SCHEME_REF)
    c.advance(0.01, OMEGA, L.SCHEME_REF, L.PREC_FP64, nsteps=1)  # capture
    c.synchronize()
    t = time.perf_counter()
    c.advance(0.01, OMEGA, L.SCHEME_REF, L.PREC_FP64, nsteps=1)
    c.synchronize()
    dt = time.perf_counter() - t
    c.close()
    pairs = 4 * ((nv + nl) * nl - nl)
    return {"workload": f"icosahedral level {level}: {nv} vertices, {nl} faces; 32 Gaussian vortices (seed 1); "
                        "reference RK4 step, FP64, dt=0.01",
            "ms_per_step": 1e3 * dt, "pairs_per_step": pairs, "pairs_per_s": pairs / dt}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--compare-steps", type=int, default=10,
                    help="FP64 and FP32C steps from the same state: FP64 rate and FP32C error (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--targets-per-lane", type=int, default=0)
    ap.add_argument("--splits", type=int, default=0)
    ap.add_argument("--near-factor", type=float, default=None)
    ap.add_argument("--staging", choices=["lds", "scalar"], default=None)
    ap.add_argument("--form", choices=["precision", "ensemble", "serving"], default="precision")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-max-level", type=int, default=8)
    ap.add_argument("--ensemble-members", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU; sharding over several GPUs (target slices, per-stage "
              "all-gather of the source records) is designed but not run. Use --gpus 1.", file=sys.stderr)
        return 2

    import lpmkokkos_amd as L

    if args.form == "ensemble":
        return ensemble_form(args)
    if args.form == "serving":
        out = serving_form(args.serve_requests, args.serve_max_level)
        print(json.dumps(out), flush=True)
        return 0

    t0 = time.perf_counter()
    mesh, fx, farea, vz, fz = build_problem(args.level)
    nv, nl = mesh.nverts, len(fx)
    pairs_eval = (nv + nl) * nl - nl
    pairs_step = 4 * pairs_eval
    log(f"[bench] level {args.level}: {nv} vertices, {nl} active faces, {pairs_eval:.4e} pairs/eval "
        f"(setup {time.perf_counter() - t0:.1f} s)")

    ctx = L.Context(0)
    ctx.set_state(mesh.vx, vz, fx, fz, farea, np.zeros(nl, np.uint8))
    if args.targets_per_lane:
        ctx.set_launch(args.targets_per_lane, args.splits)
    if args.near_factor is not None:
        ctx.set_near_factor(args.near_factor)
    if args.staging:
        ctx.set_source_staging(args.staging == "lds")

    # FP32C: warmup, then K timed steps replayed from the captured graph
    ctx.solve(L.PREC_FP32C, L.SCHEME_LAGR)
    ctx.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=args.warmup)
    ctx.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps)")
    sampler = ClockSampler().start()
    t1 = time.perf_counter()
    ctx.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=args.steps)
    ctx.synchronize()
    t_fp32c = time.perf_counter() - t1
    clock = sampler.stop()
    ms_step = 1e3 * t_fp32c / args.steps
    rate = pairs_step * args.steps / t_fp32c
    clk = (f"{clock['gfx_mhz_mean']:.0f} MHz mean over the XCDs, {clock['gfx_mhz_min_xcd']}-"
           f"{clock['gfx_mhz_max_xcd']}" if "gfx_mhz_mean" in clock else clock.get("error"))
    log(f"[bench] FP32C: {ms_step:.1f} ms/step, {rate:.4e} pairs/s; GFX clock {clk}")

    # pair-kernel launch time with HIP events (one step, eager launches)
    ctx.enable_timing(True)
    ctx.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=1, use_graph=False)
    nlaunch, kernel_ms = ctx.pair_timing()
    ctx.enable_timing(False)
    achieved = FLOPS_PER_PAIR * pairs_eval / (kernel_ms * 1e-3) / 1e12
    # the measured HBM traffic belongs to the level-8 workload at the default launch shape
    default_launch = (args.level == 8 and not args.targets_per_lane and not args.splits
                      and args.near_factor is None and args.staging in (None, "lds"))
    log(f"[bench] pair kernel: {nlaunch} launches, {kernel_ms:.2f} ms each, {achieved:.2f} TFLOP/s")

    # FP64 and FP32C from the same initial state over --compare-steps steps (the FP64 run is
    # also the FP64 rate), then FP32C error against FP64: positions after the steps, velocity
    # from the step kernel and stream function at the final FP64 state
    rate64 = rate64_ref = err = binding = None
    nsteps_cmp = args.compare_steps
    if nsteps_cmp > 0:
        c64 = L.Context(0)
        c64.set_state(mesh.vx, vz, fx, fz, farea, np.zeros(nl, np.uint8))
        c64.solve(L.PREC_FP64, L.SCHEME_LAGR)  # entry velocity (stage 1 of the first step)
        c64.synchronize()
        t2 = time.perf_counter()
        c64.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP64, nsteps=nsteps_cmp)
        c64.synchronize()
        t_fp64 = time.perf_counter() - t2
        rate64 = pairs_step * nsteps_cmp / t_fp64
        log(f"[bench] FP64: {1e3 * t_fp64 / nsteps_cmp:.1f} ms/step, {rate64:.4e} pairs/s ({nsteps_cmp} steps)")
        if not args.no_accuracy:
            c32 = L.Context(0)
            c32.set_state(mesh.vx, vz, fx, fz, farea, np.zeros(nl, np.uint8))
            c32.advance(DT, OMEGA, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=nsteps_cmp)
            a, b = c64.get_state(), c32.get_state()
            x64 = np.concatenate([a["vx"], a["fx"]])
            x32 = np.concatenate([b["vx"], b["fx"]])
            dx = np.linalg.norm(x32 - x64, axis=1)  # |x| = 1: absolute = relative
            # one evaluation at the FP64 state in both precisions
            c64.solve(L.PREC_FP64, L.SCHEME_LAGR, want_psi=True)
            a = c64.get_state()
            c32.set_state(a["vx"], a["vzeta"], a["fx"], a["fzeta"], farea, np.zeros(nl, np.uint8))
            c32.solve(L.PREC_FP32C, L.SCHEME_LAGR)  # velocity with the kernel the timed steps use
            bv = c32.get_state()
            c32.solve(L.PREC_FP32C, L.SCHEME_LAGR, want_psi=True)
            bp = c32.get_state()
            c32.close()
            u64 = np.concatenate([a["vu"], a["fu"]])
            u32 = np.concatenate([bv["vu"], bv["fu"]])
            p64 = np.concatenate([a["vpsi"], a["fpsi"]])
            p32 = np.concatenate([bp["vpsi"], bp["fpsi"]])
            du = np.linalg.norm(u32 - u64, axis=1)
            umax = np.linalg.norm(u64, axis=1).max()
            err = {
                "steps": nsteps_cmp,
                "positions_max_rel": float(dx.max()),
                "positions_rms_rel": float(np.sqrt(np.mean(dx ** 2))),
                "velocity_max_rel": float(du.max() / umax),
                "velocity_rms_rel": float(np.sqrt(np.mean(du ** 2)) / np.sqrt(np.mean(np.sum(u64 ** 2, axis=1)))),
                "psi_max_rel": float(np.abs(p32 - p64).max() / np.abs(p64).max()),
                "psi_rms_rel": float(np.sqrt(np.mean((p32 - p64) ** 2) / np.mean(p64 ** 2))),
            }
            log(f"[bench] FP32C vs FP64: {err}")
        # the reference scheme (literal 1 - x.y, no renormalisation) from the FP64 state
        c64.advance(DT, OMEGA, L.SCHEME_REF, L.PREC_FP64, nsteps=1)  # warm (capture + entry velocity)
        c64.synchronize()
        t3 = time.perf_counter()
        c64.advance(DT, OMEGA, L.SCHEME_REF, L.PREC_FP64, nsteps=1)
        c64.synchronize()
        t_ref = time.perf_counter() - t3
        rate64_ref = pairs_step / t_ref
        log(f"[bench] FP64 reference scheme: {1e3 * t_ref:.1f} ms/step, {rate64_ref:.4e} pairs/s")
        binding = drop_in_binding(c64, farea, nl, pairs_step, t_ref)
        log(f"[bench] copy-in/copy-out binding: {binding['ms_per_step']:.1f} ms/step "
            f"({binding['copy_ms_per_step']:.1f} ms of copies), {binding['pairs_per_s']:.4e} pairs/s")
        c64.close()
    ctx.close()

    c1 = None
    if not args.no_c1:
        c1 = c1_gpu_step()
        log(f"[bench] level-7 FP64 reference RK4 step: {c1['ms_per_step']:.1f} ms, {c1['pairs_per_s']:.4e} pairs/s")

    ensemble = None
    if not args.no_ensemble:
        ensemble = ensemble_run(args.ensemble_members)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        log(f"[bench] CPU baseline: {cpu['value']:.4e} pairs/s on {cpu['cores']} threads ({cpu['cpu_model']})")

    serving = None
    if args.serve_requests > 0:
        # the serving loop runs with its own default launch shapes and grids, as
        # `python -m lpmkokkos_amd.serve` does: forget the process-wide rules the
        # batching configuration tuned for its sizes
        L._lib.call("lpm_shape_rules_clear")
        L._lib.call("lpm_ens_rules_clear")
        serving = serving_form(args.serve_requests, args.serve_max_level)
        log(f"[bench] serving sample: p50 {serving['latency_p50_s']:.2f} s, p99 {serving['latency_p99_s']:.2f} s, "
            f"{serving['pairs_per_s']:.4e} pairs/s")

    out = {
        "metric": "Biot-Savart pair-interactions/s, FP32c vs FP64 at 1.31M src; p50/p99 latency",
        "value": rate,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"icosahedral level {args.level}: {nv} vertex + {nl} face targets, {nl} face sources; "
                        "32 Gaussian vortices (seed 1, sigma 0.1) + Omega=2pi; Lagrangian RK4 dt=0.005",
            "pairs_per_step": pairs_step,
            "evaluations_per_step": 4,
        },
        "rk4_steps_per_s": 1e3 / ms_step,
        "gpu_clock_during_fp32c": clock,
        "fp64_pairs_per_s": rate64,
        "fp64_reference_scheme_pairs_per_s": rate64_ref,
        "drop_in_binding_pcie_inclusive": binding,
        "fp32c_vs_fp64_speedup": rate / rate64 if rate64 else None,
        "fp32c_vs_fp64_speedup_bound": FP32C_VS_FP64_ISSUE_BOUND,
        "fp32c_vs_fp64_speedup_bound_note": "instruction-issue limit of the two pair loops: FP64 FMA issues at half "
                                            "the FP32 rate, 45.0 vs 22.2 SIMD cycles per 64 pairs",
        "c1_fp64_reference_step": c1,
        "ensemble": ensemble,
        "fp32c_error_vs_fp64": err,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_VECTOR_PEAK_TFLOPS,
            # the same against the peak scaled to the clock the box ran the timed steps at
            "frac_at_measured_clock": (achieved / (FP32_VECTOR_PEAK_TFLOPS * clock["gfx_mhz_mean"] / 2400.0)
                                       if "gfx_mhz_mean" in clock else None),
            "traffic": TRAFFIC_BYTES_PER_LAUNCH if default_launch else None,
            "traffic_source": "profiles/k_pair_f32c_lds_level8_counters.md" if default_launch else None,
            "kernel": "k_pair_f32c",
            "kernel_ms": kernel_ms,
            "flops_per_pair": FLOPS_PER_PAIR,
            "pairs_per_launch": pairs_eval,
        },
        "cpu_baseline": cpu,
        "serving": serving,
        "latency_p50_s": serving["latency_p50_s"] if serving else None,
        "latency_p99_s": serving["latency_p99_s"] if serving else None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
