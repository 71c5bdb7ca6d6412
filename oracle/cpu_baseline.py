"""CPU baseline of the BVE direct sum on the host's cores.

TEST INFRASTRUCTURE: the timed loop is the oracle's reference-structured
restatement (per target, a sequential FP64 sweep over every tree face with
the mask and the literal 1 - x.y, src/LpmBVEKernels.hpp:163-186, 285-305; the
RK4 step of src/LpmBVERK4_Impl.hpp:42-130).  bench.py runs this file as a
child process, so that its OpenMP thread binding (OMP_PROC_BIND=close,
OMP_PLACES=cores, one thread per physical core) applies to the oracle only and
never pins the threads of the process that drives the GPU.  It prints one JSON
line.

Workloads (the same inputs the GPU measurement uses):
  C0  level 4 (2,562 vertices, 6,820 tree faces), Rossby-Haurwitz wave n=4,
      w = K = 0.1, Omega = 2 pi: 100 reference RK4 steps of dt = 0.01, in full
      (one initial evaluation + 4 per step).
  C2  level 8, 32 Gaussian vortices (seed 1, amplitudes N(0,1), sigma 0.1,
      area-weighted mean removed): one evaluation at every 64th vertex and
      every 64th tree face, scaled to all targets (extrapolated).
  C1  level 7, the same field: one evaluation over all targets, in full, when
      the rate measured on C2 predicts at most `--c1-budget` seconds; otherwise
      a strided subset sized to that budget (reported as extrapolated).

Rates: reference-loop iterations/s ((targets) x (tree faces)) and algorithmic
pair interactions/s ((vertices + leaves) x leaves - leaves per evaluation).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, the folder itself comes first on the path; the package is wanted
sys.path = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402


def cpu_info() -> dict:
    """CPU model, logical CPUs this process may use, and the physical cores
    among them (distinct (package, core) pairs)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    return {"cpu_model": model, "logical_cpus": len(cpus), "physical_cores": len(cores)}


def gauss_state(m, seed=1, nvort=32, sigma=0.1):
    c, a = O.gauss_params(seed, nvort, 1.0)
    vz = O.ic_eval(m.vx, centers=c, amps=a, sigma=sigma)
    fz = O.ic_eval(m.fx, centers=c, amps=a, sigma=sigma)
    lv = m.fmask == 0
    mu = float(np.sum(fz[lv] * m.farea[lv]) / np.sum(m.farea[lv]))
    return vz - mu, fz - mu


def alg_pairs(m) -> int:
    nl = int((m.fmask == 0).sum())
    return (m.nv + nl) * nl - nl


def one_eval(m, fz, stride: int) -> tuple[float, int]:
    """One evaluation (vertex and face velocity) at every stride-th target;
    returns (seconds, targets evaluated)."""
    t0 = time.perf_counter()
    O.velocity(m.vx, m.fx, fz, m.farea, m.fmask, False, stride=stride)
    O.velocity(m.fx, m.fx, fz, m.farea, m.fmask, True, stride=stride)
    dt = time.perf_counter() - t0
    return dt, len(range(0, m.nv, stride)) + len(range(0, m.nf, stride))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--c1-budget", type=float, default=45.0, help="seconds allowed for the C1 evaluation")
    ap.add_argument("--c2-stride", type=int, default=64)
    args = ap.parse_args()
    # before the OpenMP runtime loads (it reads these variables once, and binds
    # this thread): one thread per physical core this process may use, at most
    # the OMP_NUM_THREADS it was given
    info = cpu_info()
    threads = max(1, min(info["physical_cores"], int(os.environ.get("OMP_NUM_THREADS", "0") or 16)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    O.lib()
    out = {"unit": "pair-interactions/s", "kind": "port", "cores": threads,
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}
    out.update(info)

    # C0 in full: 100 reference RK4 steps (401 evaluations over the tree)
    m0 = O.Mesh(4)
    rh = dict(rh_n=4, rh_w=0.1, rh_K=0.1)
    st = O.State(m0, O.ic_eval(m0.vx, **rh), O.ic_eval(m0.fx, **rh))
    t0 = time.perf_counter()
    st.init_velocity(0)
    for _ in range(100):
        st.step(0.01, 2 * np.pi, 0)
    t_c0 = time.perf_counter() - t0
    evals = 401
    out["C0"] = {"workload": "level 4 RH wave, 100 reference RK4 steps dt=0.01 (in full, 401 evaluations)",
                 "seconds": t_c0, "rk4_steps_per_s": 100 / t_c0,
                 "pairs_per_s": evals * alg_pairs(m0) / t_c0,
                 "iterations_per_s": evals * (m0.nv + m0.nf) * m0.nf / t_c0}
    print(f"[cpu] C0: {t_c0:.2f} s", file=sys.stderr, flush=True)

    # C2: one evaluation on a strided subset, scaled to all targets
    m2 = O.Mesh(8)
    _, fz2 = gauss_state(m2)
    s = args.c2_stride
    dt2, n2 = one_eval(m2, fz2, s)
    nt2 = m2.nv + m2.nf
    frac = n2 / nt2
    it_rate = n2 * m2.nf / dt2
    pair_rate = frac * alg_pairs(m2) / dt2
    out["C2"] = {"workload": f"level 8, 32 Gaussian vortices: one evaluation at every {s}th target "
                             f"({n2} of {nt2}), scaled to all targets (extrapolated)",
                 "seconds": dt2, "pairs_per_s": pair_rate, "iterations_per_s": it_rate,
                 "extrapolated_eval_s": dt2 / frac, "extrapolated_rk4_step_s": 4 * dt2 / frac}
    print(f"[cpu] C2 subset: {dt2:.2f} s", file=sys.stderr, flush=True)

    # C1: in full if the C2 rate predicts it fits the budget
    m1 = O.Mesh(7)
    _, fz1 = gauss_state(m1)
    nt1 = m1.nv + m1.nf
    predicted = nt1 * m1.nf / it_rate
    s1 = 1 if predicted <= args.c1_budget else int(np.ceil(predicted / (0.67 * args.c1_budget)))
    dt1, n1 = one_eval(m1, fz1, s1)
    frac1 = n1 / nt1
    out["C1"] = {"workload": "level 7, 32 Gaussian vortices: one evaluation over all targets" if s1 == 1 else
                 f"level 7, 32 Gaussian vortices: one evaluation at every {s1}th target, scaled (extrapolated)",
                 "seconds": dt1, "pairs_per_s": frac1 * alg_pairs(m1) / dt1,
                 "iterations_per_s": n1 * m1.nf / dt1, "eval_s": dt1 / frac1,
                 "rk4_step_s": 4 * dt1 / frac1, "rk4_step_note": "4 evaluations per RK4 step"}
    print(f"[cpu] C1: {dt1:.2f} s", file=sys.stderr, flush=True)

    out["value"] = pair_rate
    out["sample"] = (f"C2 (level 8): one FP64 evaluation of the reference loop at every {s}th of {nt2} targets "
                     f"x {m2.nf} tree-face sources, {dt2:.1f} s, scaled; reference loop iterations/s = {it_rate:.4g}; "
                     f"{out['cores']} OpenMP threads bound one per physical core")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
