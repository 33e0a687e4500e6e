"""Autotuner of the batched form: leaf size, sort tile and traversal launch shape.

The search space:
  bucket                 leaf size (cells with <= bucket bodies are leaves): 4, 8, 16.
                         Leaves are never approximated (nbody.wgsl:71-84), so a
                         larger bucket only replaces approximated cells by exact
                         direct terms: every candidate is at least as accurate as
                         the reference default 4 (nbody/src/lib.rs:9).
  sort_tile              keys per radix-sort workgroup: 2048, 1024
  traverse_block         traversal workgroup size: 256, 128, 64 threads
  traverse_waves_per_cu  persistent traversal wavefronts per CU: 32, 24, 20, 16, 12
  streams                groups of simulations stepped on streams of their own, so
                         their kernels overlap: 1, 2
  topo_tile              sorted positions per topology workgroup: 512, 1024

The launch-shape choices leave every result bit-identical; the leaf size defines
the tree, and the CPU restatement reproduces the tree and forces for any leaf size.
Candidates are timed on the workload itself (whole HIP-graph steps after a
warm-up) by coordinate descent from the defaults (traversal waves and groups as
one joint coordinate; passes repeated while one improves, at most three), and
the winner is cached per
workload signature (sizes histogram, softening, theta, device) in a JSON file,
so later runs of the same ensemble shape reuse it without measuring.
"""
from __future__ import annotations

import itertools
import json
import math
import os
import time
from dataclasses import asdict, dataclass, replace

import numpy as np

from .batch import Batch
from .bh import Config, NbodyGpu

SPACE = {
    "bucket": (4, 8, 16),
    "sort_tile": (2048, 1024),
    "traverse_block": (256, 128, 64),
    "traverse_waves_per_cu": (32, 24, 20, 16, 12),
    "streams": (1, 2),
    "topo_tile": (512, 1024),
}

# Coordinates of the descent.  Resident traversal waves and the number of
# groups are searched together: two groups overlap well only with fewer
# traversal waves per CU, so neither improves alone from the one-group optimum.
COORDINATES = (("bucket",), ("sort_tile",), ("topo_tile",), ("traverse_block",),
               ("traverse_waves_per_cu", "streams"))

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "galaxy_amd", "tune.json")


@dataclass(frozen=True)
class Choice:
    bucket: int = 4
    sort_tile: int = 2048
    traverse_block: int = 256
    traverse_waves_per_cu: int = 32
    streams: int = 1
    topo_tile: int = 512


def signature(sizes, config: Config, dt: float, device: str) -> str:
    """Workload key: count of simulations per power-of-two size class, physics, device."""
    hist = np.bincount(np.floor(np.log2(np.asarray(sizes, np.float64))).astype(np.int64), minlength=32)
    classes = ",".join(f"{k}:{int(c)}" for k, c in enumerate(hist) if c)
    return (f"sims={len(sizes)};sizes={classes};theta={config.theta};eps={config.softening_kpc};"
            f"maxd={config.max_depth};dt={dt};dev={device}")


def measure(gpu: NbodyGpu, states, config: Config, dt: float, choice: Choice, steps: int = 5,
            warmup: int = 2) -> float:
    """Seconds per whole batched step (HIP graph) of the workload at `choice`.  The
    warm-up of two steps captures both step graphs (first step and steady state)
    outside the timed steps."""
    cfg = replace(config, bucket=choice.bucket)
    b = Batch(gpu, states, cfg, dt, use_graph=True, sort_tile=choice.sort_tile,
              traverse_block=choice.traverse_block, traverse_waves_per_cu=choice.traverse_waves_per_cu,
              streams=choice.streams, topo_tile=choice.topo_tile)
    try:
        b.step(warmup)
        t0 = time.perf_counter()
        b.step(steps)
        return (time.perf_counter() - t0) / steps
    finally:
        b.close()


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(gpu: NbodyGpu, states, config: Config, dt: float, cache: str | None = DEFAULT_CACHE,
         steps: int = 5, log=None) -> dict:
    """Best Choice for this ensemble.  Returns {"choice", "seconds_per_step",
    "measured": [(choice, s/step), ...], "cached": bool, "signature"}."""
    sizes = [len(s) for s in states]
    dev = gpu.info().get("driver_info", "")
    key = signature(sizes, config, dt, dev)
    if cache:
        hit = _load(cache).get(key)
        if hit:
            return {"choice": Choice(**hit["choice"]), "seconds_per_step": hit["seconds_per_step"],
                    "measured": hit.get("measured", []), "cached": True, "signature": key}
    best = Choice()
    seen = {}

    def cost(c: Choice) -> float:
        if c not in seen:
            seen[c] = measure(gpu, states, config, dt, c, steps=steps)
            if log:
                log(f"  tune {asdict(c)}: {seen[c] * 1e3:.3f} ms/step")
        return seen[c]

    cost(best)
    for _ in range(3):  # coordinate descent, repeated while a pass improves (timings are kept)
        start = best
        for names in COORDINATES:
            for values in itertools.product(*(SPACE[n] for n in names)):
                cand = replace(best, **dict(zip(names, values)))
                if cost(cand) < cost(best):
                    best = cand
        if best == start:
            break
    result = {"choice": best, "seconds_per_step": seen[best],
              "measured": [(asdict(c), s) for c, s in seen.items()], "cached": False, "signature": key}
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        data = _load(cache)
        data[key] = {"choice": asdict(best), "seconds_per_step": seen[best],
                     "measured": result["measured"]}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, cache)
    return result


def bodies(sizes) -> int:
    return int(math.fsum(sizes))
