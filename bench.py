#!/usr/bin/env python3
"""Benchmark: Barnes-Hut body-steps/s, batched+served vs plain at N=65,536.

One JSON line on stdout (rank 0, one GPU):
  value          batched ensemble throughput (body-steps/s): 256 rotating discs,
                 N log-uniform in [512, 65536], HIP-graph steps, at the launch shape
                 the autotuner picks (config.batched_choice; the tuner's measurements
                 in autotune; --choice or --no-tune to fix it).  batched_one_group: the
                 same shape as one group with one launch per phase.
  plain_*        the drop-in at N=65,536 (cold collapse), driven step by step through
                 the synchronous C ABI exactly like galaxy-bh-gpu-tree-parallel.rs:345-350
  large_single   one 1,048,576-body binary disc encounter (theta 0.6) through the
                 beyond-cap drop-in tree
  serving        the request stream: 20,000 Poisson arrivals ramping 50 -> 2,000 req/s,
                 served continuously by the library's worker thread; p50/p99 latency
                 from arrival to completion, body-steps/s, offered load
  roofline       the traversal kernel (dominant) of the batched step, FP32 VALU flops;
                 HBM traffic from profiles/traverse_batched256_traffic.json; the whole
                 step against the BH-model-v1 roofline with per-phase HBM rates
  cpu_baseline   the oracle's f64 BH #2A CPU trajectory (cpu_flat_evolve restatement),
                 single thread, on bounded samples of C0-C3 (headline: 16 of the C3
                 simulations), with the host's core count and model
--only plain|batched|large|copy|serving|cpu runs one leg (for profiling) and prints it.
A step = kick/drift + full tree rebuild + force traversal + final kick, inputs resident in HBM.
"""
from __future__ import annotations

import argparse
import json
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
              streams=choice.streams, topo_tile=choice.topo_tile)
    total = w.bodies
    b.step(max(warmup, 1))
    t0 = time.perf_counter()
    b.step(steps)
    el = time.perf_counter() - t0
    # per-phase kernel times (HIP events) of one un-captured step, averaged
    phases = [b.phase_ms() for _ in range(3)]
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    inter = cells = 0
    for i in range(len(w.states)):
        st, _ = b.force_evidence(i)
        inter += int(st[:, 1].sum() + st[:, 2].sum())
        cells += b.tree_counts(i)[0]
    b.close()
    one_group = None
    if choice.streams != 1:
        # One group of 256 simulations on one stream: one launch per phase (the tuned
        # headline steps two groups, each with its own launches), at the best one-group
        # shape the tuner measured (else the headline's shape with one stream).  Its
        # phase timing gives the traversal kernel alone, one launch of the whole set.
        one = replace(choice, streams=1)
        if tuning is not None:
            ones = [(t, c) for c, t in tuning["measured"] if c["streams"] == 1 and c["bucket"] == choice.bucket]
            if ones:
                one = autotune.Choice(**min(ones, key=lambda x: x[0])[1])
        b1 = Batch(gpu, w.states, cfg, w.dt, use_graph=True, sort_tile=one.sort_tile,
                   traverse_block=one.traverse_block, traverse_waves_per_cu=one.traverse_waves_per_cu,
                   streams=1, topo_tile=one.topo_tile)
        b1.step(max(warmup, 1))
        t1 = time.perf_counter()
        b1.step(steps)
        el1 = time.perf_counter() - t1
        ph1 = [b1.phase_ms() for _ in range(3)]
        b1.close()
        one_group = dict(choice=asdict(one), ms_per_step=1e3 * el1 / steps, body_steps_per_s=total * steps / el1,
                         phase_ms={k: float(np.mean([p[k] for p in ph1])) for k in ph1[0]})
    tune_out = None
    if tuning is not None:
        tune_out = {"choice": asdict(choice), "cached": tuning["cached"],
                    "measured_ms_per_step": [dict(c, ms=1e3 * t) for c, t in tuning["measured"]]}
    return dict(sims=len(w.states), bodies=total, steps=steps, seconds=el,
                body_steps_per_s=total * steps / el, ms_per_step=1e3 * el / steps, phase_ms=avg,
                interactions=inter, cells=cells, choice=asdict(choice), tuning=tune_out,
                one_group=one_group)


def serving_run(gpu, requests=20000, choice=None, threads=16):
    """The C4 stream through the server's worker thread (continuous batching).
    Request states are generated before the clock starts (the clients' inputs);
    each is submitted at its arrival time and its latency runs from that arrival
    to its completion on the server clock."""
    from concurrent.futures import ThreadPoolExecutor

    from galaxy_amd import ics
    from galaxy_amd.serve import Server

    reqs = ics.config4_requests(requests)
    phys = {k: ics.request_physics(k) for k in ("disc", "collapse")}

    def make(r):
        _, kind, n, seed, _ = r
        return ics.request_state(kind, n, seed)

    t_gen = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        states = list(ex.map(make, reqs))
    t_gen = time.perf_counter() - t_gen
    log(f"  generated {len(reqs)} request states ({sum(len(s) for s in states)} bodies) in {t_gen:.1f} s")
    body_steps = sum(n * k for _, _, n, _, k in reqs)
    shape = {k: choice[k] for k in ("sort_tile", "traverse_block", "traverse_waves_per_cu", "streams",
                                                 "topo_tile")
             if k in choice} if choice else {}
    srv = Server(gpu, hbm_budget_bytes=200 * 10**9, steps_per_round=10, **shape)
    srv.start()
    t0 = srv.now()
    idx_of, done, lag = {}, {}, []
    max_q = max_infl = 0
    max_res = 0
    i, last_log, last_stat = 0, 0.0, 0.0
    while len(done) < len(reqs):
        now = srv.now() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            _, kind, n, _, steps = reqs[i]
            cfg, dt = phys[kind]
            rid = srv.submit(states[i], cfg, dt, steps)
            lag.append(srv.now() - t0 - reqs[i][0])
            states[i] = None
            idx_of[rid] = i
            i += 1
        if now - last_stat >= 0.5:
            q, infl, res = srv.stats()
            max_q, max_infl, max_res = max(max_q, q), max(max_infl, infl), max(max_res, res)
            last_stat = now
        if now - last_log >= 10.0:
            log(f"  serving t={now:6.1f} s: {i} submitted, {len(done)} done, in flight {infl}, queued {q}")
            last_log = now
        wait = 0.05 if i >= len(reqs) else min(0.05, max(0.0, reqs[i][0] - now))
        for c in srv.poll(wait=wait):
            done[c.id] = c
            srv.take(c.id, keep=False)
    srv.stop()
    srv.close()
    arrive = {rid: t0 + reqs[k][0] for rid, k in idx_of.items()}
    lat = np.array([c.finish_time - arrive[rid] for rid, c in done.items()])
    wait_q = np.array([c.admit_time - arrive[rid] for rid, c in done.items()])
    makespan = max(c.finish_time for c in done.values()) - t0
    span = reqs[-1][0]
    small = np.array([reqs[idx_of[rid]][4] == 10 for rid in done])
    return dict(
        workload="Poisson stream of 20,000 'advance k steps' requests" if requests == 20000
        else f"Poisson stream of {requests} 'advance k steps' requests",
        requests=len(reqs), arrival_span_s=span, rate_req_per_s=[50.0, 2000.0],
        n_range=[1024, 262144], k_choices=[10, 100, 1000], hbm_budget_bytes=200 * 10**9,
        steps_per_round=10, ics="alternating rotating disc (C0 physics, bucket 1) and cold collapse "
                                "(C1 physics, bucket 4), seed 10^6 + index",
        state_generation_s=t_gen, seconds=makespan, body_steps=body_steps,
        body_steps_per_s=body_steps / makespan,
        offered_body_steps_per_s=body_steps / span,
        offered_over_sustained=(body_steps / span) / (body_steps / makespan),
        p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
        p50_latency_k10_s=float(np.percentile(lat[small], 50)) if small.any() else None,
        p50_queue_wait_s=float(np.percentile(wait_q, 50)),
        max_submit_lag_s=float(max(lag)), max_inflight=max_infl, max_queued=max_q,
        max_resident_bytes=max_res, failed=int(sum(c.status != 0 for c in done.values())))


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "model_name": model}


def cpu_baseline():
    """The reference CPU path (cpu_flat_evolve, galaxy-bh-gpu-tree-parallel.rs:124-148:
    f64 BH #2A tree rebuilt every step, flat traversal, KDK; single-threaded) as the
    oracle's restatement, on bounded samples of each workload.  The four samples run
    at once, one host thread each (each is the single-threaded reference path)."""
    from concurrent.futures import ThreadPoolExecutor

    from galaxy_amd import ics
    from oracle import cpu_ref

    def leg(name, states, config, dt, steps, label):
        cfg = cpu_ref.make_config(theta=config.theta, softening=config.softening_kpc, g=config.g,
                                  bucket=config.bucket, max_depth=config.max_depth)
        t0 = time.perf_counter()
        bs = 0
        for st in states:
            s = st.astype(np.float64)
            cpu_ref.flat_evolve(s[:, 0], s[:, 1], s[:, 4], s[:, 5], s[:, 2], steps, dt, cfg)
            bs += len(s) * steps
        el = time.perf_counter() - t0
        return name, dict(value=bs / el, unit="body-steps/s", seconds=el, body_steps=bs, sample=label)

    c0, c1, c2, c3 = ics.config0(), ics.config1(), ics.config2(), ics.config3()
    pick = list(range(0, 256, 16))  # 16 of the 256 simulations, every sixteenth
    c3_states = [c3.states[i] for i in pick]
    legs = [
        ("c0", c0.states, c0.config, c0.dt, c0.steps, "C0 full run: 768 bodies x 200 steps"),
        ("c1", c1.states, c1.config, c1.dt, 10, "C1 prefix: 65,536 bodies, first 10 of 1,000 steps "
                                                 "(rate extrapolated to the run)"),
        ("c2", c2.states, c2.config, c2.dt, 2, "C2 prefix: 1,048,576 bodies, first 2 of 500 steps "
                                                "(rate extrapolated to the run)"),
        ("c3", c3_states, c3.config, c3.dt, 10,
         f"C3 sample: simulations {pick[0]}, {pick[1]}, ..., {pick[-1]} of 256 "
         f"({sum(len(s) for s in c3_states)} bodies), first 10 of 500 steps"),
    ]
    with ThreadPoolExecutor(len(legs)) as ex:
        out = dict(ex.map(lambda a: leg(*a), legs))
    head = out["c3"]
    return dict(value=head["value"], unit="body-steps/s", cores=1, kind="port",
                sample=head["sample"] + "; cpu_flat_evolve restated (oracle), f64 BH #2A, one thread",
                host=host_cpu(), legs=out,
                note="each leg is the single-threaded reference path on one host thread; the four legs "
                     "ran concurrently on separate cores")


def bh_model_v1(bodies, cells, interactions, ms_per_step, phase_ms):
    """The BH-model-v1 roofline of one step: build bytes 112 N + 64 C, KDK bytes
    64 N at 8 TB/s, traversal flops 18 I at 157.3 TFLOP/s; fraction = T_roof / T.
    Per phase, the model's bytes over the phase's kernel time against 8 TB/s."""
    n, c, i = bodies, cells, interactions
    build_b, kdk_b, flops = 112 * n + 64 * c, 64 * n, FLOPS_PER_INTERACTION * i
    t_roof = (build_b + kdk_b) / (HBM_PEAK_GBS * 1e9) + flops / (FP32_PEAK_TFLOPS * 1e12)
    phases = {
        # (model bytes, kernels timed)
        "kdk_bounds": (64 * n + 8 * n, ["kick_drift_bounds"]),
        "keys": (16 * n, ["keys"]),
        "sort": (64 * n + 8 * n, ["radix_sort"]),
        "topology_aggregates": (16 * n + 64 * c, ["topology", "top_aggregates"]),
    }
    per = {}
    for name, (byts, keys) in phases.items():
        ms = sum(phase_ms[k] for k in keys)
        gbs = byts / (ms * 1e-3) / 1e9
        per[name] = {"model_bytes": int(byts), "ms": ms, "GBps": gbs, "frac_of_8TBps": gbs / HBM_PEAK_GBS}
    return {"bodies": n, "cells": c, "interactions": i, "build_bytes": int(build_b), "kdk_bytes": int(kdk_b),
            "traversal_flops": int(flops), "t_roof_ms": 1e3 * t_roof, "t_measured_ms": ms_per_step,
            "fraction": 1e3 * t_roof / ms_per_step, "phase_hbm": per}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=20000)
    ap.add_argument("--no-tune", action="store_true", help="batched leg at the default launch shape and leaf size 4")
    ap.add_argument("--tune-cache", default=None, help="JSON cache of autotuner results (default: none, tune every run)")
    ap.add_argument("--choice", default=None,
                    help="fixed batched choice instead of tuning, e.g. bucket=4,sort_tile=2048,"
                         "traverse_block=256,traverse_waves_per_cu=16,streams=2")
    ap.add_argument("--no-large", action="store_true", help="skip the 1,048,576-body single simulation")
    ap.add_argument("--only", choices=["plain", "batched", "large", "copy", "serving", "cpu"], default=None,
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
    if args.only == "cpu":
        print(json.dumps(cpu_baseline()), flush=True)
        return

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
    if args.only == "serving":
        from dataclasses import asdict

        from galaxy_amd.autotune import Choice

        c = asdict(choice or Choice(traverse_waves_per_cu=16, streams=2))
        print(json.dumps(serving_run(gpu, requests=args.serving_requests, choice=c)), flush=True)
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
        log(f"serving: C4 stream of {args.serving_requests} requests ...")
        serving = serving_run(gpu, requests=args.serving_requests, choice=batched["choice"])
        log(f"  serving {serving['body_steps_per_s']:.3e} body-steps/s, p50 {serving['p50_latency_s']:.2f} s "
            f"p99 {serving['p99_latency_s']:.2f} s")
    gpu.close()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (four single-threaded legs at once) ...")
        cpu = cpu_baseline()
        log(f"  cpu {cpu['value']:.3e} body-steps/s on the C3 sample")

    trav_ms = batched["phase_ms"]["traversal"]
    traffic = measured_traffic()
    achieved = FLOPS_PER_INTERACTION * batched["interactions"] / (trav_ms * 1e-3) / 1e12
    plain_trav = FLOPS_PER_INTERACTION * plain["interactions"] / (plain["traversal_ms"] * 1e-3) / 1e12
    model = bh_model_v1(batched["bodies"], batched["cells"], batched["interactions"], batched["ms_per_step"],
                        batched["phase_ms"])
    one_launch = None
    if batched["one_group"]:
        # the kernel alone: one traversal launch over all 256 simulations (one group)
        ms1 = batched["one_group"]["phase_ms"]["traversal"]
        a1 = FLOPS_PER_INTERACTION * batched["interactions"] / (ms1 * 1e-3) / 1e12
        one_launch = dict(achieved=a1, frac=a1 / FP32_PEAK_TFLOPS, launch_ms=ms1,
                          choice=batched["one_group"]["choice"])
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
                        f"autotuned leaf size {batched['choice']['bucket']}, steps {args.warmup + 1}-"
                        f"{args.warmup + args.steps} of the 500-step run; plain drop-in: cold collapse N=65536, "
                        "theta 0.5, eps 0.005, dt 5e-4, bucket 4; serving: the C4 stream of "
                        f"{args.serving_requests} requests",
            "sims": batched["sims"],
            "bodies": batched["bodies"],
            "theta": 0.5,
            "batched_choice": batched["choice"],
            "launches_per_phase": int(batched["choice"].get("streams", 1)),
        },
        "autotune": batched["tuning"],
        "plain_body_steps_per_s": plain["body_steps_per_s"],
        "plain_ms_per_step": plain["ms_per_step"],
        "batched_vs_plain": batched["body_steps_per_s"] / plain["body_steps_per_s"],
        "batched_one_group": batched["one_group"],
        "batched_one_group_vs_plain": (batched["one_group"]["body_steps_per_s"] / plain["body_steps_per_s"]
                                       if batched["one_group"] else None),
        "serving": serving,
        "served_vs_plain": serving["body_steps_per_s"] / plain["body_steps_per_s"] if serving else None,
        "large_single": large,
        "phase_ms": batched["phase_ms"],
        "roofline": {
            "bound": "valu",
            "compute": "fp32 vector ALU (no MFMA: the traversal is not a dense contraction)",
            "kernel": "k_traverse_nodes",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "frac_of_unpacked_fp32_78_6": achieved / 78.6,
            "traffic": traffic["traffic_bytes"] if traffic else None,
            "traffic_source": "profiles/traverse_batched256_traffic.json (rocprofv3 FETCH_SIZE x%.3f + "
                              "WRITE_SIZE x%.3f per launch, summed over the step's launches)"
                              % (traffic["fetch_scale"], traffic["write_scale"]) if traffic else None,
            "interactions_per_step": batched["interactions"],
            # one traversal launch per group of simulations (config.batched_choice streams);
            # kernel_ms_per_step sums the groups' launches, timed one after the other
            "launches_per_step": int(batched["choice"].get("streams", 1)),
            "kernel_ms_per_step": trav_ms,
            "one_launch": one_launch,
            "plain_traversal_tflops": plain_trav,
            "plain_traversal_ms": plain["traversal_ms"],
            "step_bh_model_v1": model,
        },
        "cpu_baseline": cpu,
        "plain": plain,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
