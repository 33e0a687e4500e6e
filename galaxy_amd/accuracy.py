"""FP64 accuracy and conservation of the Barnes-Hut step on every workload.

For each workload the report gives:
  * the force error of the step's f32 Barnes-Hut accelerations against f64 direct
    summation on probe bodies (probe_error, nbody/src/lib.rs:370-395: RMS and max
    relative error over probes k*(N-1)/(P-1)), at the first and the last force
    state of the run;
  * the relative drift of the total energy (kinetic + Plummer-softened pair
    potential, f64) and the drift of Lz over the run.  A run that starts with
    Lz = 0 (cold collapse) reports |dLz| divided by sum(m |x| |v|) at the end.

Workloads (seeded, as the benchmark builds them):
  c0  rotating disc N=768, 200 steps, drop-in
  c1  cold collapse N=65,536, 1,000 steps, drop-in
  c2  binary disc encounter N=1,048,576, 500 steps, beyond-cap drop-in
  c3  256 rotating discs, N log-uniform in [512, 65,536], 500 steps, batched
  c4  the first 200 requests of the serving stream (discs and collapses, N
      log-uniform in [1,024, 262,144], k in {10, 100, 1,000}), run through the server

The f64 direct sums and the O(N^2) potential run on the GPU (gbh_diag_direct,
gbh_diag_energy); both equal the CPU restatement bit for bit (tests/test_gpu_diag.py).

    python -m galaxy_amd.accuracy [--configs c0,c1,c2,c3,c4] [--steps-scale 1.0] [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import ics
from .batch import Batch
from .bh import Config, NbodyGpu, ParallelTreeRuntime
from .serve import Server


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lz_scale(state: np.ndarray) -> float:
    s = state.astype(np.float64)
    return float(np.sum(s[:, 2] * np.hypot(s[:, 0], s[:, 1]) * np.hypot(s[:, 4], s[:, 5])))


def drift(e0: dict, e1: dict, final_state: np.ndarray) -> dict:
    de = (e1["energy"] - e0["energy"]) / abs(e0["energy"]) if e0["energy"] != 0 else float("nan")
    dlz = e1["lz"] - e0["lz"]
    if abs(e0["lz"]) > 0:
        rel_lz, lz_basis = dlz / abs(e0["lz"]), "|Lz0|"
    else:
        scale = lz_scale(final_state)
        rel_lz, lz_basis = (dlz / scale if scale > 0 else float("nan")), "sum m|x||v| (Lz0 = 0)"
    return {"energy0": e0["energy"], "energy1": e1["energy"], "rel_energy_drift": de,
            "lz0": e0["lz"], "lz1": e1["lz"], "rel_lz_drift": rel_lz, "lz_basis": lz_basis}


def dropin(gpu, state, cfg: Config, dt: float, steps: int, probes: int, large: bool) -> dict:
    rt = ParallelTreeRuntime(gpu)
    ev = gpu.create_evolving_state(state)
    tree = rt.create_tree(gpu, ev, large=large)
    e0 = gpu.energy(ev, cfg.softening_kpc, cfg.g)
    rt.rebuild(gpu, ev, tree, cfg)
    rt.force(gpu, ev, tree, cfg)
    p0 = gpu.probe_error(ev, cfg, gpu.read_evolving_accelerations(ev), probes)
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.kick_drift(ev, dt)
        rt.rebuild(gpu, ev, tree, cfg)
        rt.force(gpu, ev, tree, cfg)
        gpu.final_kick(ev, dt)
    run_s = time.perf_counter() - t0
    p1 = gpu.probe_error(ev, cfg, gpu.read_evolving_accelerations(ev), probes)
    e1 = gpu.energy(ev, cfg.softening_kpc, cfg.g)
    final = gpu.read_evolving_state(ev)
    out = {"n": len(state), "steps": steps, "run_seconds": run_s, "probe_first": p0, "probe_last": p1}
    out.update(drift(e0, e1, final))
    tree.close()
    ev.close()
    rt.close()
    return out


def state_metrics(gpu, state0, state1, accel1, cfg: Config, probes: int) -> dict:
    ev0 = gpu.create_evolving_state(state0)
    e0 = gpu.energy(ev0, cfg.softening_kpc, cfg.g)
    ev0.close()
    ev1 = gpu.create_evolving_state(state1)
    e1 = gpu.energy(ev1, cfg.softening_kpc, cfg.g)
    p1 = gpu.probe_error(ev1, cfg, accel1, probes)
    ev1.close()
    out = {"n": len(state0), "probe_last": p1}
    out.update(drift(e0, e1, state1))
    return out


def summarize(rows: list, key) -> dict:
    v = np.array([key(r) for r in rows], np.float64)
    a = np.abs(v)
    return {"median_abs": float(np.median(a)), "max_abs": float(a.max()), "count": int(len(v))}


def run_c0(gpu, scale):
    w = ics.config0()
    return dropin(gpu, w.states[0], w.config, w.dt, max(1, int(w.steps * scale)), len(w.states[0]), False)


def run_c1(gpu, scale):
    w = ics.config1()
    return dropin(gpu, w.states[0], w.config, w.dt, max(1, int(w.steps * scale)), 4096, False)


def run_c2(gpu, scale):
    w = ics.config2()
    return dropin(gpu, w.states[0], w.config, w.dt, max(1, int(w.steps * scale)), 4096, True)


def run_c3(gpu, scale, count=256):
    w = ics.config3(count=count)
    steps = max(1, int(w.steps * scale))
    b = Batch(gpu, w.states, w.config, w.dt, use_graph=True)
    b.step(steps)
    rows = []
    for i, s0 in enumerate(w.states):
        s1, a1 = b.read(i)
        rows.append(state_metrics(gpu, s0, s1, a1, w.config, min(len(s0), 256)))
    b.close()
    return {"sims": len(rows), "bodies": w.bodies, "steps": steps,
            "probe_last_rms": summarize(rows, lambda r: r["probe_last"]["rms_relative"]),
            "probe_last_max": summarize(rows, lambda r: r["probe_last"]["max_relative"]),
            "rel_energy_drift": summarize(rows, lambda r: r["rel_energy_drift"]),
            "rel_lz_drift": summarize(rows, lambda r: r["rel_lz_drift"])}


def run_c4(gpu, scale, requests=200):
    """The first `requests` requests of the C4 stream, submitted at once and
    served to completion (their accuracy does not depend on arrival times)."""
    srv = Server(gpu, hbm_budget_bytes=200 * 10**9, steps_per_round=10)
    reqs = {}
    for _, kind, n, seed, k in ics.config4_requests()[:requests]:
        steps = max(1, int(k * scale))
        s = ics.request_state(kind, n, seed)
        cfg, dt = ics.request_physics(kind)
        rid = srv.submit(s, cfg, dt, steps)
        reqs[rid] = (s, cfg, steps)
    done = srv.drain()
    rows = []
    for c in done:
        if c.status != 0:
            raise RuntimeError(f"request {c.id} failed with status {c.status}")
        s0, cfg, steps = reqs[c.id]
        s1, a1 = srv.take(c.id)
        r = state_metrics(gpu, s0, s1, a1, cfg, min(len(s0), 256))
        r["steps"] = steps
        rows.append(r)
    srv.close()
    return {"requests": len(rows), "body_steps": int(sum(r["n"] * r["steps"] for r in rows)),
            "probe_last_rms": summarize(rows, lambda r: r["probe_last"]["rms_relative"]),
            "probe_last_max": summarize(rows, lambda r: r["probe_last"]["max_relative"]),
            "rel_energy_drift": summarize(rows, lambda r: r["rel_energy_drift"]),
            "rel_lz_drift": summarize(rows, lambda r: r["rel_lz_drift"])}


RUNNERS = {"c0": run_c0, "c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="c0,c1,c2,c3,c4")
    ap.add_argument("--steps-scale", type=float, default=1.0, help="fraction of each workload's steps")
    ap.add_argument("--json", default=None)
    args = ap.parse_args()
    gpu = NbodyGpu()
    report = {"steps_scale": args.steps_scale}
    for name in args.configs.split(","):
        log(f"accuracy {name} ...")
        t0 = time.perf_counter()
        report[name] = RUNNERS[name](gpu, args.steps_scale)
        report[name]["wall_seconds"] = time.perf_counter() - t0
        log(f"  {name}: {json.dumps(report[name])[:400]}")
        if args.json:
            with open(args.json, "w") as f:
                json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
