#!/usr/bin/env python3
"""Benchmark: Barnes-Hut body-steps/s, batched+served vs plain at N=65,536.

One JSON line on stdout (rank 0, one GPU):
  value          batched ensemble throughput (body-steps/s): 256 rotating discs,
                 N log-uniform in [512, 65536], one launch per phase, HIP-graph steps,
                 at the launch shape the autotuner picks (config.batched_choice; the
                 tuner's measurements in autotune; --choice or --no-tune to fix it)
  plain_*        the drop-in at N=65,536 (cold collapse), driven step by step through
                 the synchronous C ABI exactly like galaxy-bh-gpu-tree-parallel.rs:345-350
  large_single   one 1,048,576-body binary disc encounter (theta 0.6) through the
                 beyond-cap drop-in tree
  serving        a bounded sample of the request stream: p50/p99 latency, body-steps/s
  roofline       the traversal kernel (dominant) of the batched step, FP32 flops; HBM
                 traffic from profiles/traverse_batched256_traffic.json
  cpu_baseline   the oracle's f64 BH #2A CPU trajectory (cpu_flat_evolve restatement),
                 single thread, on a bounded prefix of the N=65,536 run
--only plain|batched|large|copy runs one leg (for profiling) and prints its result.
A step = kick/drift + full tree rebuild + force traversal + final kick, inputs resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 (vector and matrix)
HBM_PEAK_GBS = 8000.0
FLOPS_PER_INTERACTION = 18  # delta 2, r2 4, rsqrt 4, r^-3 2, G m r^-3 2, accumulate 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per traversal launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by the copy-kernel calibration), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "traverse_batched256_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def interactions(rt_or_batch, *args):
    st, _ = rt_or_batch(*args)
    return int(st[:, 1].sum() + st[:, 2].sum())


def plain_run(gpu, steps, warmup):
    from galaxy_amd import ParallelTreeRuntime, ics

    w = ics.config1()
    rt = ParallelTreeRuntime(gpu)
    ev = gpu.create_evolving_state(w.states[0])
    tree = rt.create_tree(gpu, ev)
    cfg, dt = w.config, w.dt
    rt.rebuild(gpu, ev, tree, cfg)
    rt.force(gpu, ev, tree, cfg)

    def step():
        gpu.kick_drift(ev, dt)
        rt.rebuild(gpu, ev, tree, cfg)
        rt.force(gpu, ev, tree, cfg)
        gpu.final_kick(ev, dt)

    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    force_ms = []
    for _ in range(steps):
        step()
        force_ms.append(rt.last_force_ms(tree))
    el = time.perf_counter() - t0
    stats, _ = rt.force_evidence(ev, tree, cfg)
    inter = int(stats[:, 1].sum() + stats[:, 2].sum())
    active, leaves, maxd = rt.counts(tree)
    n = len(w.states[0])
    return dict(n=n, steps=steps, seconds=el, body_steps_per_s=n * steps / el, ms_per_step=1e3 * el / steps,
                traversal_ms=float(np.mean(force_ms)), interactions=inter, cells=active, leaves=leaves,
                max_depth=maxd)


def large_run(gpu, steps, warmup):
    """One simulation beyond the reference cap: binary disc encounter, N = 1,048,576,
    theta 0.6, through the same synchronous drop-in calls with the beyond-cap tree."""
    from galaxy_amd import ParallelTreeRuntime, ics

    w = ics.config2()
    rt = ParallelTreeRuntime(gpu)
    ev = gpu.create_evolving_state(w.states[0])
    tree = rt.create_tree(gpu, ev, large=True)
    cfg, dt = w.config, w.dt
    rt.rebuild(gpu, ev, tree, cfg)
    rt.force(gpu, ev, tree, cfg)

    def step():
        gpu.kick_drift(ev, dt)
        rt.rebuild(gpu, ev, tree, cfg)
        rt.force(gpu, ev, tree, cfg)
        gpu.final_kick(ev, dt)

    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    force_ms = []
    for _ in range(steps):
        step()
        force_ms.append(rt.last_force_ms(tree))
    el = time.perf_counter() - t0
    n = len(w.states[0])
    out = dict(n=n, steps=steps, seconds=el, body_steps_per_s=n * steps / el, ms_per_step=1e3 * el / steps,
               traversal_ms=float(np.mean(force_ms)), theta=cfg.theta)
    tree.close()
    ev.close()
    return out


def batched_run(gpu, steps, warmup, count=256, tune=True, tune_cache=None, choice=None):
    from dataclasses import asdict, replace

    from galaxy_amd import autotune, ics
    from galaxy_amd.batch import Batch

    w = ics.config3(count=count)
    tuning = None
    if choice is not None:
        tune = False
    choice = choice or autotune.Choice()
    if tune:
        log("  autotuning leaf size / sort tile / traversal launch shape ...")
        tuning = autotune.tune(gpu, w.states, w.config, w.dt, cache=tune_cache, steps=20, log=log)
        choice = tuning["choice"]
        log(f"  tuned: {asdict(choice)}")
    cfg = replace(w.config, bucket=choice.bucket)
    b = Batch(gpu, w.states, cfg, w.dt, use_graph=True, sort_tile=choice.sort_tile,
              traverse_block=choice.traverse_block, traverse_waves_per_cu=choice.traverse_waves_per_cu,
              streams=choice.streams)
    total = w.bodies
    b.step(max(warmup, 1))
    t0 = time.perf_counter()
    b.step(steps)
    el = time.perf_counter() - t0
    # per-phase kernel times (HIP events) of one un-captured step, averaged
    phases = [b.phase_ms() for _ in range(3)]
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    inter = 0
    for i in range(len(w.states)):
        st, _ = b.force_evidence(i)
        inter += int(st[:, 1].sum() + st[:, 2].sum())
    b.close()
    tune_out = None
    if tuning is not None:
        tune_out = {"choice": asdict(choice), "cached": tuning["cached"],
                    "measured_ms_per_step": [dict(c, ms=1e3 * t) for c, t in tuning["measured"]]}
    return dict(sims=len(w.states), bodies=total, steps=steps, seconds=el,
                body_steps_per_s=total * steps / el, ms_per_step=1e3 * el / steps, phase_ms=avg,
                interactions=inter, choice=asdict(choice), tuning=tune_out)


def serving_run(gpu, requests=300, rate_lo=50.0, rate_hi=2000.0, seed=777, choice=None):
    """Bounded sample of the serving stream: Poisson arrivals ramping rate_lo -> rate_hi.
    Requests carry their own physics (leaf size 4); the server uses the tuned launch shape."""
    from galaxy_amd import Config, ics
    from galaxy_amd.serve import Server

    rng = np.random.default_rng(seed)
    span = requests / ((rate_lo + rate_hi) / 2.0)
    t, arrivals = 0.0, []
    while len(arrivals) < requests:
        rate = rate_lo + (rate_hi - rate_lo) * min(t / span, 1.0)
        t += rng.exponential(1.0 / rate)
        arrivals.append(t)
    reqs = []
    for k, at in enumerate(arrivals):
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(262144)))))
        steps = int(rng.choice([10, 100, 1000]))
        if k % 2 == 0:
            s = ics.rotating_disc(n, 10**6 + k, softening=0.01)
            cfg, dt = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16), 1e-3
        else:
            s = ics.uniform_collapse(n, 10**6 + k)
            cfg, dt = Config(theta=0.5, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16), 5e-4
        reqs.append((at, s, cfg, dt, steps))
    shape = {k: choice[k] for k in ("sort_tile", "traverse_block", "traverse_waves_per_cu", "streams")
             if k in choice} if choice else {}
    srv = Server(gpu, hbm_budget_bytes=200 * 10**9, steps_per_round=10, **shape)
    t0 = srv.now()
    i, done = 0, []
    body_steps = sum(len(s) * st for _, s, _, _, st in reqs)
    while len(done) < len(reqs):
        now = srv.now() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            _, s, cfg, dt, steps = reqs[i]
            srv.submit(s, cfg, dt, steps)
            i += 1
        q, infl, _ = srv.stats()
        if q == 0 and infl == 0:
            time.sleep(max(0.0, min(0.01, reqs[i][0] - now)))
            continue
        done.extend(srv.round())
        if len(done) % 50 == 0:
            log(f"  serving: {len(done)}/{len(reqs)} done")
    el = srv.now() - t0
    lat = np.array([d.latency for d in done])
    srv.close()
    return dict(requests=len(reqs), seconds=el, body_steps=body_steps, body_steps_per_s=body_steps / el,
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                failed=int(sum(d.status != 0 for d in done)))


def cpu_baseline(steps=40):
    """Oracle (restated reference CPU path, f64 BH #2A cpu_flat_evolve), single thread."""
    from galaxy_amd import ics
    from oracle import cpu_ref

    w = ics.config1()
    s = w.states[0].astype(np.float64)
    cfg = cpu_ref.make_config(theta=w.config.theta, softening=w.config.softening_kpc, g=w.config.g,
                              bucket=w.config.bucket, max_depth=w.config.max_depth)
    t0 = time.perf_counter()
    cpu_ref.flat_evolve(s[:, 0], s[:, 1], s[:, 4], s[:, 5], s[:, 2], steps, w.dt, cfg)
    el = time.perf_counter() - t0
    n = len(s)
    return dict(value=n * steps / el, unit="body-steps/s", cores=1, kind="port",
                sample=f"cold collapse N={n}: initial force + {steps} KDK steps of cpu_flat_evolve "
                       f"(f64 BH #2A, rebuild every step), {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=300)
    ap.add_argument("--no-tune", action="store_true", help="batched leg at the default launch shape and leaf size 4")
    ap.add_argument("--tune-cache", default=None, help="JSON cache of autotuner results (default: none, tune every run)")
    ap.add_argument("--choice", default=None,
                    help="fixed batched choice instead of tuning, e.g. bucket=4,sort_tile=2048,"
                         "traverse_block=256,traverse_waves_per_cu=16")
    ap.add_argument("--no-large", action="store_true", help="skip the 1,048,576-body single simulation")
    ap.add_argument("--only", choices=["plain", "batched", "large", "copy"], default=None,
                    help="run one leg only (for profiling); prints that leg's result")
    args = ap.parse_args()
    choice = None
    if args.choice:
        from galaxy_amd.autotune import Choice

        choice = Choice(**{k: int(v) for k, v in (kv.split("=") for kv in args.choice.split(","))})
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the simulations of this workload are independent and "
              f"run on a single MI355X; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from galaxy_amd import NbodyGpu

    gpu = NbodyGpu()
    log(f"device: {gpu.info()}")
    if args.only == "copy":
        import ctypes as C

        from galaxy_amd import _lib as L

        ms = C.c_double()
        nbytes = 1 << 30
        for _ in range(3):
            if L.lib().gbh_copy_calibration(gpu._h, nbytes, C.byref(ms)) != 0:
                raise RuntimeError(L.lib().gbh_last_error().decode())
        print(json.dumps({"copy_bytes": nbytes, "copy_ms": ms.value,
                          "copy_GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)
        return
    if args.only == "plain":
        print(json.dumps(plain_run(gpu, args.steps, args.warmup)), flush=True)
        return
    if args.only == "large":
        print(json.dumps(large_run(gpu, args.steps, args.warmup)), flush=True)
        return
    if args.only == "batched":
        print(json.dumps(batched_run(gpu, args.steps, args.warmup, tune=not args.no_tune,
                                     tune_cache=args.tune_cache, choice=choice)), flush=True)
        return
    log("plain drop-in, cold collapse N=65536 ...")
    plain = plain_run(gpu, args.steps, args.warmup)
    log(f"  plain {plain['body_steps_per_s']:.3e} body-steps/s, {plain['ms_per_step']:.3f} ms/step")
    log("batched ensemble, 256 discs ...")
    batched = batched_run(gpu, args.steps, args.warmup, tune=not args.no_tune, tune_cache=args.tune_cache,
                          choice=choice)
    log(f"  batched {batched['body_steps_per_s']:.3e} body-steps/s, {batched['ms_per_step']:.3f} ms/step")
    large = None
    if not args.no_large:
        log("large single simulation, binary disc N=1048576 ...")
        large = large_run(gpu, args.steps, args.warmup)
        log(f"  large {large['body_steps_per_s']:.3e} body-steps/s, {large['ms_per_step']:.3f} ms/step")
    serving = None
    if not args.no_serving:
        log("serving sample ...")
        serving = serving_run(gpu, requests=args.serving_requests, choice=batched["choice"])
        log(f"  serving p50 {serving['p50_latency_s']:.3f} s p99 {serving['p99_latency_s']:.3f} s")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline()

    trav_ms = batched["phase_ms"]["traversal"]
    traffic = measured_traffic()
    achieved = FLOPS_PER_INTERACTION * batched["interactions"] / (trav_ms * 1e-3) / 1e12
    plain_trav = FLOPS_PER_INTERACTION * plain["interactions"] / (plain["traversal_ms"] * 1e-3) / 1e12
    line = {
        "metric": "body-steps/s batched+served vs plain at N=65,536; p50/p99 request latency",
        "value": batched["body_steps_per_s"],
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded discs / cold collapse)",
        "config": {
            "workload": "batched ensemble of 256 rotating exponential discs, N log-uniform in [512, 65536] "
                        f"({batched['bodies']} bodies), theta 0.5, eps 0.01, dt 1e-3, "
                        f"autotuned leaf size {batched['choice']['bucket']}; "
                        "plain drop-in: cold collapse N=65536, theta 0.5, eps 0.005, dt 5e-4, bucket 4",
            "sims": batched["sims"],
            "bodies": batched["bodies"],
            "theta": 0.5,
            "batched_choice": batched["choice"],
        },
        "autotune": batched["tuning"],
        "plain_body_steps_per_s": plain["body_steps_per_s"],
        "plain_ms_per_step": plain["ms_per_step"],
        "batched_vs_plain": batched["body_steps_per_s"] / plain["body_steps_per_s"],
        "serving": serving,
        "large_single": large,
        "phase_ms": batched["phase_ms"],
        "roofline": {
            "bound": "mfma",
            "compute": "fp32 vector ALU (no MFMA: the traversal is not a dense contraction)",
            "kernel": "k_traverse_nodes",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic["traffic_bytes"] if traffic else None,
            "traffic_source": "profiles/traverse_batched256_traffic.json (rocprofv3 FETCH_SIZE x%.3f + "
                              "WRITE_SIZE x%.3f per launch)" % (traffic["fetch_scale"], traffic["write_scale"])
                              if traffic else None,
            "interactions_per_launch": batched["interactions"],
            # a step's traversal: one launch per group of simulations (config.batched_choice
            # streams); kernel_ms sums the groups' launches, timed one after the other
            "launches_per_step": int(batched["choice"].get("streams", 1)),
            "kernel_ms": trav_ms,
            "plain_traversal_tflops": plain_trav,
            "plain_traversal_ms": plain["traversal_ms"],
        },
        "cpu_baseline": cpu,
        "plain": plain,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
