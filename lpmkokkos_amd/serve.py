"""Serving form: a continuous-batching request loop for BVE forecasts.

Requests (mesh level, initial field, number of RK4 steps) arrive by a seeded
Poisson process.  Between steps the loop admits every request that has
arrived while a slot and the device-memory budget allow, keeps each admitted
request's particle state resident in HBM (one context per request), advances
all of them together -- one ensemble step is four persistent pair kernels over
the flattened work of every active request -- and retires requests that have
completed their steps, freeing their memory for the next arrivals.  A new
request's state is built on the device: positions copied from the per-level
mesh cached in device memory, the initial field evaluated by a kernel, so
admission needs no host mesh, no host field evaluation and no upload.

Each request's result is bit-identical to running it alone: a member keeps
the decomposition it has alone and writes only its own partial slots.

Requests below level 8 are batched in lanes by size class; level-8 requests
run in a large lane of their own (see Server), so short requests are not held
to a level-8 step.

Checkpoint and resume: run(..., stop_after_steps=k, checkpoint_dir=d) stops
admitting after k batched steps, lets the large lanes finish, writes every
batched request's state at its step boundary to a netCDF checkpoint
(Context.checkpoint) and the queue to d/manifest.json; load_manifest(d)
returns the requests for a new server, whose results are those the first
server would have produced, bit for bit.

Reported: completion latency (finish - arrival) p50 / p99 and aggregate
pair interactions per second over the run.

    python -m lpmkokkos_amd.serve --requests 2000 --budget-gb 200
    python -m lpmkokkos_amd.serve --requests 200 --stop-after-steps 500 --checkpoint-dir ck
    python -m lpmkokkos_amd.serve --resume ck
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PREC_FP32C, SCHEME_LAGR, SEED_CUBED_SPHERE
from .bve import Context, VorticityInitialCondition
from .ensemble import Ensemble, MemberSpec, make_member, rng_draws, state_bytes

LEVEL_PROBS = {4: 0.30, 5: 0.30, 6: 0.20, 7: 0.15, 8: 0.05}


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    spec: MemberSpec
    steps: int = 20
    # filled by the server
    ctx: Context | None = None
    bytes: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    steps_done: int = 0
    admit_s: float = 0.0  # wall time of building the request's state
    pairs: float = 0.0
    result: dict | None = field(default=None, repr=False)
    checkpoint: str | None = None  # resume from this file (steps_done steps already taken)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def make_ic_zero() -> VorticityInitialCondition:
    return VorticityInitialCondition()


def tree_sizes(level: int, seed: int = 0) -> tuple[int, int]:
    """Vertices and leaves of a level-`level` tree: 2 + 10 4^L and 20 4^L
    (icosahedral), 2 + 6 4^L and 6 4^L (cubed sphere)."""
    if seed == SEED_CUBED_SPHERE:
        return 2 + 6 * 4 ** level, 6 * 4 ** level
    return 2 + 10 * 4 ** level, 20 * 4 ** level


def pairs_per_step(level: int, seed: int = 0) -> float:
    """Pair interactions of one RK4 step (four evaluations, self pairs excluded)."""
    nv, nl = tree_sizes(level, seed)
    return 4.0 * ((nv + nl) * nl - nl)


def make_requests(n: int, rate: float, seed: int = 4, level_probs=None, steps: int = 20) -> list[Request]:
    """n requests with exponential inter-arrival times of mean 1/rate; level
    drawn from level_probs; RH wavenumber uniform on 2..6 and 4..16 Gaussian
    vortices (amplitude N(0, 0.05), sigma 0.15), all from `seed`."""
    probs = level_probs or LEVEL_PROBS
    levels = sorted(probs)
    cum = np.cumsum([probs[l] for l in levels])
    cum /= cum[-1]
    u, _ = rng_draws(seed, 5 * n)
    t = 0.0
    reqs = []
    for i in range(n):
        t += -math.log(1.0 - u[5 * i]) / rate
        lev = levels[int(np.searchsorted(cum, u[5 * i + 1], side="right"))] if u[5 * i + 1] < cum[-1] else levels[-1]
        spec = MemberSpec(level=lev, k_factor=0.95 + 0.1 * u[5 * i + 2], rh_n=2 + int(5 * u[5 * i + 3]),
                          nvort=4 + int(13 * u[5 * i + 4]), amp=0.05, sigma=0.15, vortex_seed=100000 * seed + i)
        reqs.append(Request(rid=i, arrival=t, spec=spec, steps=steps))
    return reqs


def load_manifest(directory: str) -> list[Request]:
    """The requests a stopped server left (run(..., checkpoint_dir=directory)):
    started ones resume from their checkpoints and arrive at once; the others
    keep their arrival offsets past the stop."""
    with open(os.path.join(directory, "manifest.json")) as f:
        doc = json.load(f)
    out = []
    for e in doc["requests"]:
        ck = e["checkpoint"]
        out.append(Request(rid=e["rid"], arrival=e["arrival"], spec=MemberSpec(**e["spec"]), steps=e["steps"],
                           steps_done=e["steps_done"],
                           checkpoint=os.path.join(directory, ck) if ck else None))
    return out


def request_bytes(spec: MemberSpec) -> int:
    return state_bytes(*tree_sizes(spec.level, spec.seed))


class _Lane:
    """A batched lane: one ensemble stepped one RK4 step at a time."""

    def __init__(self, device: int, lo: int, hi: int):
        self.lo, self.hi = lo, hi  # levels served, inclusive
        self.ens = Ensemble(device)
        self.stream = self.ens.stream()
        self.members: list[Request] = []
        self.inflight = False


class Server:
    """Continuous batching in lanes.  Requests below `large_level` are batched
    by size class: each class (levels up to each of `batch_levels`) has its own
    lane -- one ensemble, one persistent pair kernel per evaluation over all of
    its members' work, stepped one RK4 step at a time so that membership can
    change between steps -- so a small request is held to the steps of
    requests of its own size, not to a level-7 step.  A request at
    `large_level` or above runs in a large lane (one of `large_lanes` streams,
    all its steps queued at once).  One large lane is the default: large
    requests then run one after another, so the first of two finishes in about
    half the time it takes when both share the device.  The GPU interleaves the
    lanes' workgroups.

    Every lane's stream is created once, at construction: the runtime shares
    its hardware queues (GPU_MAX_HW_QUEUES, 4 unless set before the first HIP
    call) among all streams and runs a shared queue's work in order, so lanes
    that share a queue wait for each other; `python -m lpmkokkos_amd.serve`
    and bench.py ask for 8.  A large request that finds every large lane busy
    waits for one; requests behind it are admitted."""

    def __init__(self, device: int = 0, budget_bytes: float = 200e9, max_active: int = 256, dt: float = 0.005,
                 omega: float = 2 * math.pi, prec: int = PREC_FP32C, keep_results: bool = False,
                 log=None, large_level: int = 8, large_lanes: int = 1, batch_levels=(5, 6, 7)):
        self.device = device
        self.budget = budget_bytes
        self.max_active = max_active
        self.dt, self.omega, self.prec = dt, omega, prec
        self.keep_results = keep_results
        if large_lanes < 0:
            raise ValueError(f"large_lanes must be >= 0, got {large_lanes}")
        self.large_level = large_level
        bounds = sorted({b for b in batch_levels if b < large_level - 1} | {large_level - 1})
        self.lanes = [_Lane(device, lo, hi) for lo, hi in zip([0] + [b + 1 for b in bounds[:-1]], bounds)]
        # large lanes: empty ensembles used for their streams
        self.lane_owners = [Ensemble(device) for _ in range(large_lanes)]
        self.lane_streams = [e.stream() for e in self.lane_owners]
        self.lane_of: dict[int, int] = {}  # request id -> large lane
        self.lanes_max_busy = 0
        self.large: list[Request] = []  # requests in large lanes
        # retired contexts not yet destroyed: freeing device memory waits for
        # the whole GPU, so it is done when nothing is in flight (or when an
        # admission needs the memory)
        self.retired: list[Context] = []
        self.used = 0      # bytes of live requests
        self.retired_bytes = 0
        self.log = log or (lambda s: None)
        queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        if len(self.lanes) + large_lanes + 1 > queues:
            self.log(f"[serve] {len(self.lanes) + large_lanes} lane streams share {queues} hardware queues: "
                     "lanes may wait for each other (set GPU_MAX_HW_QUEUES before the first HIP call)")

    @property
    def batch(self) -> list[Request]:
        return [r for lane in self.lanes for r in lane.members]

    @property
    def active(self) -> list[Request]:
        return self.batch + self.large

    def lane_for(self, level: int) -> _Lane:
        return next(lane for lane in self.lanes if lane.lo <= level <= lane.hi)

    def warm(self, meshes):
        """Build the device mesh cache for the (seed, level) pairs `meshes`
        (once per process)."""
        for seed, lev in sorted(set(meshes)):
            c = Context(self.device, stream=self.lanes[0].stream)
            c.set_state_level(lev, make_ic_zero(), seed_kind=seed)
            c.close()

    def free_lane(self) -> int | None:
        busy = set(self.lane_of.values())
        return next((k for k in range(len(self.lane_streams)) if k not in busy), None)

    def _admit(self, r: Request, now: float):
        large = r.spec.level >= self.large_level
        t = time.perf_counter()
        # a batched request queues on its lane's ensemble stream, a large
        # request on a free large lane's
        if large:
            k = self.free_lane()
            self.lane_of[r.rid] = k
            stream = self.lane_streams[k]
        else:
            lane = self.lane_for(r.spec.level)
            stream = lane.stream
        if r.checkpoint:
            r.ctx = Context(self.device, stream=stream)
            r.ctx.restore(r.checkpoint)
        else:
            r.ctx = make_member(r.spec, self.device, stream=stream)
        r.admit_s = time.perf_counter() - t
        r.bytes = request_bytes(r.spec)
        self.used += r.bytes
        r.admitted = now
        if large:
            r.ctx.advance(self.dt, self.omega, SCHEME_LAGR, self.prec, nsteps=r.steps - r.steps_done)
            self.large.append(r)
            self.lanes_max_busy = max(self.lanes_max_busy, len(self.large))
        else:
            lane.ens.add(r.ctx)
            lane.members.append(r)

    def _release(self, r: Request):
        if r.spec.level >= self.large_level:
            self.large.remove(r)
            del self.lane_of[r.rid]
        else:
            lane = self.lane_for(r.spec.level)
            lane.ens.remove(r.ctx)
            lane.members.remove(r)
        self.retired.append(r.ctx)
        r.ctx = None
        self.used -= r.bytes
        self.retired_bytes += r.bytes

    def _retire(self, r: Request, now: float):
        if self.keep_results:
            r.result = r.ctx.get_state()
        else:
            r.ctx.synchronize()
        r.finished = now
        self._release(r)

    def _free_retired(self):
        for c in self.retired:
            c.close()
        self.retired.clear()
        self.retired_bytes = 0

    def _suspend(self, pending: list[Request], now: float, directory: str):
        """Checkpoint every batched request at its step boundary, release it,
        and write the queue (with the pending requests) to the manifest."""
        os.makedirs(directory, exist_ok=True)
        entries = []
        for r in list(self.batch):
            name = f"request_{r.rid}.nc"
            r.ctx.checkpoint(os.path.join(directory, name), time=r.steps_done * self.dt, omega=self.omega,
                             level=r.spec.level)
            self._release(r)
            entries.append((r, name, 0.0))
        entries += [(r, None, max(r.arrival - now, 0.0)) for r in pending]
        doc = {"dt": self.dt, "omega": self.omega, "prec": self.prec,
               "requests": [{"rid": r.rid, "arrival": arr, "spec": dataclasses.asdict(r.spec), "steps": r.steps,
                             "steps_done": r.steps_done if name else 0, "checkpoint": name}
                            for r, name, arr in entries]}
        with open(os.path.join(directory, "manifest.json"), "w") as f:
            json.dump(doc, f, indent=1)
        return len(entries)

    def run(self, requests: list[Request], stop_after_steps: int | None = None,
            checkpoint_dir: str | None = None) -> dict:
        """Serve `requests` to completion, or, with stop_after_steps, stop after
        that many batched steps (summed over the lanes): admit no more, let
        running steps and large lanes finish, and checkpoint what is left into
        checkpoint_dir (see load_manifest)."""
        if stop_after_steps is not None and not checkpoint_dir:
            raise ValueError("stop_after_steps needs a checkpoint_dir")
        suspended = 0
        for r in requests:
            if request_bytes(r.spec) > self.budget:
                raise ValueError(f"request {r.rid} (level {r.spec.level}) exceeds the memory budget")
            if r.spec.level >= self.large_level and not self.lane_streams:
                raise ValueError(f"request {r.rid} (level {r.spec.level}) needs a large lane and the server has none "
                                 f"(large_lanes=0, large_level={self.large_level})")
        self.warm((r.spec.seed, r.spec.level) for r in requests)
        pending = sorted(requests, key=lambda r: r.arrival)
        done: list[Request] = []
        t0 = time.perf_counter()
        busy = 0.0
        pairs = 0.0
        steps = 0
        last_log = last = t0
        while pending or self.batch or self.large:
            tn = time.perf_counter()
            inflight = any(lane.inflight for lane in self.lanes)
            if inflight or self.large:
                busy += tn - last
            last = tn
            now = tn - t0
            # completions
            for lane in self.lanes:
                if lane.inflight and lane.ens.done():
                    lane.inflight = False
                    steps += 1
                    for r in list(lane.members):
                        r.steps_done += 1
                        inc = pairs_per_step(r.spec.level, r.spec.seed)
                        r.pairs += inc
                        pairs += inc
                        if r.steps_done >= r.steps:
                            self._retire(r, now)
                            done.append(r)
            for r in list(self.large):
                if r.ctx.done():
                    inc = pairs_per_step(r.spec.level, r.spec.seed) * (r.steps - r.steps_done)
                    r.steps_done = r.steps
                    r.pairs += inc
                    pairs += inc
                    self._retire(r, now)
                    done.append(r)
            inflight = any(lane.inflight for lane in self.lanes)
            if self.retired and not inflight and not self.large:
                self._free_retired()
            if stop_after_steps is not None and steps >= stop_after_steps:
                if not inflight and not self.large:
                    suspended = self._suspend(pending, now, checkpoint_dir)
                    pending = []
                    self._free_retired()
                    break
                time.sleep(0.0005)
                continue
            # admission in arrival order while a slot and memory allow; a lane
            # changes membership only between its steps, and a large request
            # needs a free large lane (later arrivals may pass a waiting one)
            i = 0
            while i < len(pending) and pending[i].arrival <= now and len(self.active) < self.max_active:
                r = pending[i]
                if r.spec.level >= self.large_level:
                    if self.free_lane() is None:
                        i += 1
                        continue
                elif self.lane_for(r.spec.level).inflight:
                    i += 1
                    continue
                need = request_bytes(r.spec)
                if self.used + self.retired_bytes + need > self.budget:
                    if self.retired and self.used + need <= self.budget and not inflight:
                        self._free_retired()  # memory pressure: free now, even while a large lane runs
                    else:
                        break
                self._admit(pending.pop(i), now)
            for lane in self.lanes:
                if not lane.inflight and lane.members:
                    lane.ens.advance(self.dt, self.omega, SCHEME_LAGR, self.prec, nsteps=1)
                    lane.inflight = True
            inflight = any(lane.inflight for lane in self.lanes)
            if tn - last_log > 10.0:
                last_log = tn
                self.log(f"[serve] t={now:.0f}s done={len(done)} batched={len(self.batch)} large={len(self.large)} "
                         f"pending={len(pending)} batched steps={steps} memory={self.used / 1e9:.1f} GB")
            if not inflight and not self.large:
                if pending:
                    time.sleep(min(max(pending[0].arrival - (time.perf_counter() - t0), 0.0), 0.05))
            else:
                time.sleep(0.0005)
        wall = time.perf_counter() - t0
        self._free_retired()
        lat = np.array([r.latency for r in done] or [0.0])
        below = [r.latency for r in done if r.spec.level < self.large_level] or [0.0]
        return {
            "requests": len(done),
            "suspended": suspended,
            "wall_s": wall,
            "device_busy_s": busy,
            "batched_steps": steps,
            "batch_lanes": [[lane.lo, lane.hi] for lane in self.lanes],
            "large_lane_requests": int(sum(1 for r in done if r.spec.level >= self.large_level)),
            "large_lanes": len(self.lane_streams),
            "large_lanes_max_busy": self.lanes_max_busy,
            "pairs": pairs,
            "pairs_per_s": pairs / wall,
            "pairs_per_busy_s": pairs / busy if busy else None,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()),
            "latency_p50_s_below_large": float(np.percentile(below, 50)),
            "latency_p50_s_by_level": {int(l): float(np.percentile([r.latency for r in done if r.spec.level == l], 50))
                                       for l in sorted({r.spec.level for r in done})},
            "latency_p99_s_by_level": {int(l): float(np.percentile([r.latency for r in done if r.spec.level == l], 99))
                                       for l in sorted({r.spec.level for r in done})},
            "latency_max_s": float(lat.max()),
            "admission_mean_s": float(np.mean([r.admit_s for r in done] or [0.0])),
            "admission_max_s": float(np.max([r.admit_s for r in done] or [0.0])),
            "levels": {int(l): int(sum(1 for r in done if r.spec.level == l)) for l in sorted({r.spec.level for r in done})},
        }

    def close(self):
        for lane in self.lanes:
            lane.ens.close()
        for e in self.lane_owners:
            e.close()


def measure_service_times(levels, steps: int = 20, device: int = 0) -> dict:
    """Device seconds for one request of each level, run alone (capacity
    calibration)."""
    out = {}
    for lev in levels:
        c = make_member(MemberSpec(level=lev, vortex_seed=7), device)
        c.advance(0.005, 2 * math.pi, SCHEME_LAGR, PREC_FP32C, nsteps=1)
        c.synchronize()
        n = 2 if lev >= 8 else 4
        t = time.perf_counter()
        c.advance(0.005, 2 * math.pi, SCHEME_LAGR, PREC_FP32C, nsteps=n)
        c.synchronize()
        out[lev] = (time.perf_counter() - t) / n * steps
        c.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--load", type=float, default=0.8, help="arrival rate as a fraction of measured capacity")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-level", type=int, default=8)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--stop-after-steps", type=int, default=None,
                    help="drain after this many batched steps into --checkpoint-dir")
    ap.add_argument("--checkpoint-dir", default=None)
    ap.add_argument("--resume", default=None, metavar="DIR",
                    help="serve the requests a drained server left in DIR (its manifest.json)")
    ap.add_argument("--large-lanes", type=int, default=1,
                    help="lanes for level-8 requests (one runs them one after another, two side by side)")
    args = ap.parse_args(argv)
    if args.stop_after_steps is not None and not args.checkpoint_dir:
        ap.error("--stop-after-steps needs --checkpoint-dir")
    # one hardware queue per lane stream (read at the first HIP call, below;
    # the runtime's default is 4)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    probs = {l: p for l, p in LEVEL_PROBS.items() if l <= args.max_level}
    tot = sum(probs.values())
    probs = {l: p / tot for l, p in probs.items()}
    log = lambda s: print(s, file=sys.stderr, flush=True)
    if args.resume:
        reqs = load_manifest(args.resume)
        with open(os.path.join(args.resume, "manifest.json")) as f:
            saved = json.load(f)
        srv = Server(budget_bytes=args.budget_gb * 1e9, log=log, dt=saved["dt"], omega=saved["omega"],
                     prec=saved["prec"], large_lanes=args.large_lanes)
        res = srv.run(reqs, stop_after_steps=args.stop_after_steps, checkpoint_dir=args.checkpoint_dir)
        srv.close()
        res.update({"resumed_from": args.resume, "budget_gb": args.budget_gb})
        print(json.dumps(res))
        return 0
    svc = measure_service_times(sorted(probs), steps=args.steps)
    log(f"[serve] service seconds per request by level: {svc}")
    mean_service = sum(probs[l] * svc[l] for l in probs)
    rate = args.load / mean_service
    reqs = make_requests(args.requests, rate, seed=args.seed, level_probs=probs, steps=args.steps)
    srv = Server(budget_bytes=args.budget_gb * 1e9, log=log, large_lanes=args.large_lanes)
    res = srv.run(reqs, stop_after_steps=args.stop_after_steps, checkpoint_dir=args.checkpoint_dir)
    srv.close()
    res.update({"arrival_rate_per_s": rate, "service_s_by_level": svc, "load": args.load,
                "level_probs": probs, "budget_gb": args.budget_gb})
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
