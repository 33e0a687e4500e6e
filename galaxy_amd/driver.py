"""BH #2D driver: the step loop, evidence and receipt of galaxy-bh-gpu-tree-parallel.

Restates runtime/src/bin/galaxy-bh-gpu-tree-parallel.rs:280-608 on this build's
drop-in API: the same arguments, range checks and presets, the initial build and
force, `steps` kick/drift + rebuild + force + final-kick steps with no host
readback inside the loop, strict (Morton code, body index) order checks, the
FNV-mix64 tree checksum of the initial, final and same-state repeat trees, the
direct-probe force gate, the complete-rebuild benchmark (upper median) and a
receipt with schema galaxy.barnes-hut-parallel-gpu-tree-receipt.v1.

Up to --oracle-limit bodies it also runs the reference's two GPU oracles
(galaxy-bh-gpu-tree-parallel.rs:422-474): the same trajectory through the BH #2C
correctness-first GPU tree (run_serial_oracle, :225-246) and through host-built
BH #2A trees fed to the traversal (BH #2B2, run_host_oracle, :248-271), with the
reference's state and force gates; and up to 4,096 bodies the BH #2C rebuild
benchmark beside the BH #2D one (:484-505).

Differences, each recorded in the receipt itself:
  * the CPU BH #2A f64 force and trajectory comparisons (:387-420) carry status
    "skipped-test-oracle": the CPU f64 restatement is the test oracle, and tests/
    compare this path with it bit for bit and attach those sections;
  * parallel_tree_buffer_bytes keeps the reference's sparse-layout formula
    (17 N cells of 64 bytes, nbody_parallel.rs:990-992) for compatibility;
    device_tree_buffer_bytes is what this build allocates;
  * the CLI caps on steps (256) and probes (64) are lifted, and --large selects the
    beyond-cap tree (up to 2^23 bodies).

    python -m galaxy_amd.driver --preset disc --particles 512 --steps 3 --receipt out.json
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import ics
from .bh import BhError, Config, NbodyGpu, ParallelTreeRuntime, host_flat_tree

RECEIPT_SCHEMA = "galaxy.barnes-hut-parallel-gpu-tree-receipt.v1"
TREE_ORDERING = "stable-lsd-radix-4bit-morton-code; resident body index retained for equal Morton keys"
TREE_LAYOUT = "sparse-level-order-slot=depth*N+group_start"
BENCHMARK_SAMPLE_SCOPE = ("complete rebuild call including host-side uniform/bind-group setup, "
                          "command encoding, submit and synchronization")
G_KPC3_PER_MSUN_MYR2 = 4.498502151575286e-12   # nbody/src/lib.rs:6
DEFAULT_SOFTENING_KPC = 0.05                     # nbody/src/lib.rs:8
ORACLE_SKIP = {"status": "skipped-test-oracle",
               "reason": "the CPU BH #2A restatement is test infrastructure; tests/ compare this path "
                         "with it bit for bit (keys, permutation, cells, interaction lists, forces, "
                         "trajectories)"}


def checksum_words(words: np.ndarray, h: int = 0xCBF29CE484222325) -> int:
    """FNV-mix64 (galaxy-bh-gpu-tree-parallel.rs:145-149) over 32-bit words in order."""
    w = np.ascontiguousarray(words, np.uint32)
    out = C.c_uint64()
    rc = L.lib().gbh_fnv_mix64(C.c_uint64(h), w.ctypes.data_as(C.c_void_p), C.c_uint64(w.size), C.byref(out))
    if rc != 0:
        raise BhError(L.lib().gbh_last_error().decode())
    return int(out.value)


def parallel_tree_checksum(ev) -> int:
    """parallel_tree_checksum (galaxy-bh-gpu-tree-parallel.rs:150-187): meta words,
    bounds bits, entry words, then every active cell's 16 words, in order."""
    words = np.concatenate([ev.meta.view(np.uint32).ravel(), ev.entries.ravel(), ev.cells.ravel()])
    return checksum_words(words)


def validate_parallel_order(ev):
    """validate_parallel_order (galaxy-bh-gpu-tree-parallel.rs:189-203)."""
    e = ev.entries.astype(np.uint64)
    key = (e[:, 0] << np.uint64(32)) | e[:, 1]
    bad = np.flatnonzero(key[1:] <= key[:-1])
    if bad.size:
        a, b = ev.entries[bad[0], :2], ev.entries[bad[0] + 1, :2]
        raise BhError(f"BH #2D radix order is not strict by (Morton code, body index): {list(a)} then {list(b)}")


def state_error(approx: np.ndarray, ref: np.ndarray) -> dict:
    """state_error (galaxy-bh-gpu-tree-parallel.rs:62-90) of two (n, 8) states, as
    the receipt's error4 object (:92-99)."""
    a = np.asarray(approx, np.float64)
    b = np.asarray(ref, np.float64)
    pd = np.hypot(a[:, 0] - b[:, 0], a[:, 1] - b[:, 1])
    vd = np.hypot(a[:, 4] - b[:, 4], a[:, 5] - b[:, 5])
    pr = np.hypot(b[:, 0], b[:, 1])
    vr = np.hypot(b[:, 4], b[:, 5])
    return {"position_rms_relative_l2": float(np.sqrt(np.sum(pd * pd) / max(np.sum(pr * pr), 1e-30))),
            "position_max_relative": float(np.max(pd / np.maximum(pr, 1e-6))),
            "velocity_rms_relative_l2": float(np.sqrt(np.sum(vd * vd) / max(np.sum(vr * vr), 1e-30))),
            "velocity_max_relative": float(np.max(vd / np.maximum(vr, 1e-9)))}


def force_error(approx: np.ndarray, ref: np.ndarray):
    """force_error (galaxy-bh-gpu-tree-parallel.rs:101-114) of two (n, 4) accelerations."""
    a = np.asarray(approx, np.float64)
    b = np.asarray(ref, np.float64)
    rel = np.hypot(a[:, 0] - b[:, 0], a[:, 1] - b[:, 1]) / np.maximum(np.hypot(b[:, 0], b[:, 1]), 1e-30)
    return float(np.sqrt(np.sum(rel * rel) / len(rel))), float(np.max(rel))


def state_gate(name: str, e: dict):
    if (e["position_rms_relative_l2"] >= 0.03 or e["position_max_relative"] >= 0.30
            or e["velocity_rms_relative_l2"] >= 0.03 or e["velocity_max_relative"] >= 0.30):
        raise BhError(f"{name} parity failed: {e}")


def run_serial_oracle(gpu: NbodyGpu, initial: np.ndarray, steps: int, dt: float, config: Config):
    """run_serial_oracle (galaxy-bh-gpu-tree-parallel.rs:225-246): the trajectory on
    the BH #2C correctness-first GPU tree."""
    ev = gpu.create_evolving_state(initial)
    tree = gpu.create_gpu_tree(ev)
    gpu.rebuild_gpu_tree(ev, tree, config)
    gpu.force_from_gpu_tree(ev, tree, config)
    for _ in range(steps):
        gpu.kick_drift(ev, dt)
        gpu.rebuild_gpu_tree(ev, tree, config)
        gpu.force_from_gpu_tree(ev, tree, config)
        gpu.final_kick(ev, dt)
    out = gpu.read_evolving_state(ev), gpu.read_evolving_accelerations(ev)
    tree.close()
    ev.close()
    return out


def run_host_oracle(gpu: NbodyGpu, bodies: np.ndarray, steps: int, dt: float, config: Config):
    """run_host_oracle (galaxy-bh-gpu-tree-parallel.rs:248-271): BH #2A trees built on
    the host (the first from the f64 bodies, then from each read-back drifted
    state) and traversed on the GPU (BH #2B2)."""
    ev = gpu.create_evolving_state(ics.bodies_to_state(bodies))
    gpu.force_into(ev, *host_flat_tree(bodies[:, 0], bodies[:, 1], bodies[:, 4], config))
    for _ in range(steps):
        gpu.kick_drift(ev, dt)
        d = gpu.read_evolving_state(ev).astype(np.float64)
        gpu.force_into(ev, *host_flat_tree(d[:, 0], d[:, 1], d[:, 2], config))
        gpu.final_kick(ev, dt)
    out = gpu.read_evolving_state(ev), gpu.read_evolving_accelerations(ev)
    ev.close()
    return out


def upper_median(values):
    v = sorted(values)
    return v[len(v) // 2]


def make_bodies(preset: str, n: int, seed: int) -> np.ndarray:
    if preset == "disc":
        return ics.reference_disc(n, seed, 5.0e10, 3.0)
    if preset == "collision":
        return ics.reference_collision(n, seed)
    raise BhError("--preset must be disc or collision")


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description="BH #2D parallel GPU Barnes-Hut tree construction and evolution")
    ap.add_argument("--preset", default="disc")
    ap.add_argument("--particles", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--dt-myr", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=303)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--softening-kpc", type=float, default=DEFAULT_SOFTENING_KPC)
    ap.add_argument("--direct-probes", type=int, default=12)
    ap.add_argument("--oracle-limit", type=int, default=4096)
    ap.add_argument("--benchmark-warmup", type=int, default=1)
    ap.add_argument("--benchmark-repeats", type=int, default=5)
    ap.add_argument("--adapter", default=None)
    ap.add_argument("--allow-software", action="store_true")
    ap.add_argument("--require-hardware", action="store_true")
    ap.add_argument("--large", action="store_true", help="beyond-cap tree (no reference counterpart)")
    ap.add_argument("--receipt", default=None)
    return ap.parse_args(argv)


def check_args(a):
    """Range checks of galaxy-bh-gpu-tree-parallel.rs:281-311 (step and probe caps lifted)."""
    cap = 1 << 23 if a.large else 65536
    if not 2 <= a.particles <= cap:
        raise BhError(f"--particles must be in 2..={cap} for BH #2D")
    if a.preset == "collision" and a.particles < 4:
        raise BhError("collision preset requires at least 4 particles")
    if a.steps < 1:
        raise BhError("--steps must be at least 1")
    if not (np.isfinite(a.dt_myr) and 1e-6 <= a.dt_myr <= 1.0):
        raise BhError("--dt-myr must be finite and in [1e-6, 1]")
    if not (np.isfinite(a.theta) and 0.0 <= a.theta <= 2.0):
        raise BhError("--theta must be finite and in [0, 2]")
    if not (np.isfinite(a.softening_kpc) and 0.0 <= a.softening_kpc <= 100.0):
        raise BhError("--softening-kpc must be finite and in [0, 100]")
    if a.direct_probes < 1:
        raise BhError("--direct-probes must be at least 1")
    if not 2 <= a.oracle_limit <= 4096:
        raise BhError("--oracle-limit must be in 2..=4096")
    if a.benchmark_repeats == 0 or a.benchmark_repeats > 31 or a.benchmark_warmup > 10:
        raise BhError("benchmark warmup/repeats must be warmup<=10 and repeats in 1..=31")


def run(a, keep: dict | None = None) -> dict:
    """Runs the workload and returns the receipt; `keep` (optional) receives the
    initial state and the final state and accelerations."""
    check_args(a)
    bodies = make_bodies(a.preset, a.particles, a.seed)
    initial = ics.bodies_to_state(bodies)
    config = Config(theta=a.theta, softening_kpc=a.softening_kpc, g=G_KPC3_PER_MSUN_MYR2, bucket=4,
                    max_depth=48)
    gpu = NbodyGpu(a.adapter, a.allow_software)
    info = gpu.info()
    software = bool(info.get("software", False))
    if a.require_hardware and software:
        raise BhError("--require-hardware rejected the selected software GPU adapter")
    runtime = ParallelTreeRuntime(gpu)
    t_create = time.perf_counter()
    evolving = gpu.create_evolving_state(initial)
    state_create_s = time.perf_counter() - t_create
    tree = runtime.create_tree(gpu, evolving, large=a.large)

    wall0 = time.perf_counter()
    build = {k: 0.0 for k in ("bounds_seconds", "morton_seconds", "radix_seconds", "positions_seconds",
                              "topology_seconds", "aggregate_seconds")}
    stage = {name: {"transfer_seconds": 0.0, "dispatch_seconds": 0.0} for name in ("force", "kd", "fk")}

    def add_build(t):
        for k in build:
            build[k] += t[k]

    def add_stage(name, t):
        stage[name]["transfer_seconds"] += t["transfer_seconds"]
        stage[name]["dispatch_seconds"] += t["dispatch_seconds"]

    add_build(runtime.rebuild(gpu, evolving, tree, config))
    add_stage("force", runtime.force(gpu, evolving, tree, config))
    ev0 = runtime.read_evidence(gpu, tree)
    validate_parallel_order(ev0)
    initial_checksum = parallel_tree_checksum(ev0)
    for _ in range(a.steps):  # no host readback inside the loop (galaxy-bh-gpu-tree-parallel.rs:345-350)
        add_stage("kd", gpu.kick_drift(evolving, a.dt_myr))
        add_build(runtime.rebuild(gpu, evolving, tree, config))
        add_stage("force", runtime.force(gpu, evolving, tree, config))
        add_stage("fk", gpu.final_kick(evolving, a.dt_myr))
    t0 = time.perf_counter()
    final_state = gpu.read_evolving_state(evolving)
    state_readback_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    final_accel = gpu.read_evolving_accelerations(evolving)
    accel_readback_s = time.perf_counter() - t0
    ev1 = runtime.read_evidence(gpu, tree)
    validate_parallel_order(ev1)
    final_checksum = parallel_tree_checksum(ev1)
    runtime.rebuild(gpu, evolving, tree, config)
    ev2 = runtime.read_evidence(gpu, tree)
    validate_parallel_order(ev2)
    repeat_checksum = parallel_tree_checksum(ev2)
    if repeat_checksum != final_checksum:
        raise BhError("BH #2D unchanged-state repeat tree checksum changed")

    if keep is not None:
        keep.update(initial=initial, final_state=final_state, final_accel=final_accel)
    direct_count = min(a.direct_probes, a.particles)
    t0 = time.perf_counter()
    direct = gpu.probe_error(evolving, config, final_accel, direct_count)
    direct_s = time.perf_counter() - t0
    if direct["rms_relative"] >= 0.04 or direct["max_relative"] >= 0.30:
        raise BhError(f"BH #2D direct-force gate failed: rms={direct['rms_relative']} "
                      f"max={direct['max_relative']}")

    for _ in range(a.benchmark_warmup):
        runtime.rebuild(gpu, evolving, tree, config)
    samples = []
    for _ in range(a.benchmark_repeats):
        t0 = time.perf_counter()
        runtime.rebuild(gpu, evolving, tree, config)
        samples.append(time.perf_counter() - t0)
    parallel_median = upper_median(samples)

    # the GPU oracles (galaxy-bh-gpu-tree-parallel.rs:422-474)
    t0 = time.perf_counter()
    if a.particles <= a.oracle_limit:
        serial_state, serial_accel = run_serial_oracle(gpu, initial, a.steps, a.dt_myr, config)
        host_state, host_accel = run_host_oracle(gpu, bodies, a.steps, a.dt_myr, config)
        serial_err = state_error(final_state, serial_state)
        host_err = state_error(final_state, host_state)
        state_gate("BH #2C state", serial_err)
        state_gate("BH #2B2 state", host_err)
        s_rms, s_max = force_error(final_accel, serial_accel)
        h_rms, h_max = force_error(final_accel, host_accel)
        if s_rms >= 0.04 or s_max >= 0.30 or h_rms >= 0.04 or h_max >= 0.30:
            raise BhError("BH #2D force parity failed against BH #2C/BH #2B2")
        gpu_oracles = {
            "status": "executed",
            "bh2c_serial_gpu": {"state_error": serial_err, "force_rms_relative": s_rms,
                                "force_max_relative": s_max},
            "bh2b2_host_tree_gpu": {"state_error": host_err, "force_rms_relative": h_rms,
                                    "force_max_relative": h_max},
        }
    else:
        gpu_oracles = {"status": "skipped-particle-limit", "limit": a.oracle_limit,
                       "reason": "BH #2C serialized GPU oracle is intentionally capped at 4096 bodies"}
    gpu_oracle_s = time.perf_counter() - t0

    # the BH #2C rebuild beside the BH #2D one (galaxy-bh-gpu-tree-parallel.rs:484-505)
    if a.particles <= 4096:
        serial_tree = gpu.create_gpu_tree(evolving)
        for _ in range(a.benchmark_warmup):
            gpu.rebuild_gpu_tree(evolving, serial_tree, config)
        serial_samples = []
        for _ in range(a.benchmark_repeats):
            t0 = time.perf_counter()
            gpu.rebuild_gpu_tree(evolving, serial_tree, config)
            serial_samples.append(time.perf_counter() - t0)
        serial_tree.close()
        serial_median = upper_median(serial_samples)
        bh2c_serial = {"status": "executed", "samples_seconds": serial_samples, "median_seconds": serial_median,
                       "parallel_vs_serial_speedup": serial_median / parallel_median}
    else:
        bh2c_serial = {"status": "skipped-bh2c-cap", "limit": 4096}

    oracle_skip = dict(ORACLE_SKIP, limit=a.oracle_limit)
    build_total = sum(build.values())
    receipt = {
        "schema": RECEIPT_SCHEMA,
        "status": "complete",
        "phase": "BH-2D",
        "builder": "gpu-parallel-sparse-radix-v1",
        "measurement_class": "software-validation" if software else "hardware",
        "hardware_performance_claim_allowed": not software,
        "preset": a.preset,
        "particles": a.particles,
        "steps": a.steps,
        "dt_myr": a.dt_myr,
        "simulated_time_myr": a.dt_myr * a.steps,
        "seed": a.seed,
        "theta": a.theta,
        "softening_kpc": a.softening_kpc,
        "gpu": info,
        "host_tree_rebuilds": 0,
        "host_particle_readbacks_during_steps": 0,
        "force_solves": a.steps + 1,
        "evolution_tree_builds": a.steps + 1,
        "parallel_tree_buffer_bytes": int(L.lib().gbh_sparse_layout_bytes(a.particles)),
        "device_tree_buffer_bytes": runtime.buffer_bytes(tree),
        "tree": {
            "initial_checksum_fnv_mix64": f"{initial_checksum:016x}",
            "final_checksum_fnv_mix64": f"{final_checksum:016x}",
            "repeat_checksum_fnv_mix64": f"{repeat_checksum:016x}",
            "repeat_rebuild_matches": repeat_checksum == final_checksum,
            "active_cell_count": ev1.active_cell_count,
            "leaf_count": ev1.leaf_count,
            "max_depth": ev1.max_depth,
            "root_bounds_f32": [float(v) for v in ev1.bounds],
            "ordering": TREE_ORDERING,
            "layout": TREE_LAYOUT,
        },
        "final_force": {
            "bh2a_flat_status": oracle_skip["status"],
            "gpu_vs_bh2a_flat_rms_relative": None,
            "gpu_vs_bh2a_flat_max_relative": None,
            "direct_probe_count": direct_count,
            "direct_probe_rms_relative": direct["rms_relative"],
            "direct_probe_max_relative": direct["max_relative"],
        },
        "trajectory_vs_bh2a_flat_f64": oracle_skip,
        "gpu_oracles": gpu_oracles,
        "tree_build_benchmark": {
            "sample_scope": BENCHMARK_SAMPLE_SCOPE,
            "stage_timings_are_diagnostics": True,
            "warmup": a.benchmark_warmup,
            "repeats": a.benchmark_repeats,
            "parallel_samples_seconds": samples,
            "parallel_median_seconds": parallel_median,
            "bh2c_serial": bh2c_serial,
        },
        "timings_seconds": {
            "gpu_state_create_transfer": state_create_s,
            "parallel_bounds_total": build["bounds_seconds"],
            "parallel_morton_total": build["morton_seconds"],
            "parallel_radix_total": build["radix_seconds"],
            "parallel_positions_total": build["positions_seconds"],
            "parallel_topology_total": build["topology_seconds"],
            "parallel_aggregate_total": build["aggregate_seconds"],
            "parallel_tree_build_total": build_total,
            "gpu_force_transfer_total": stage["force"]["transfer_seconds"],
            "gpu_force_dispatch_total": stage["force"]["dispatch_seconds"],
            "gpu_kick_drift_transfer_total": stage["kd"]["transfer_seconds"],
            "gpu_kick_drift_dispatch_total": stage["kd"]["dispatch_seconds"],
            "gpu_final_kick_transfer_total": stage["fk"]["transfer_seconds"],
            "gpu_final_kick_dispatch_total": stage["fk"]["dispatch_seconds"],
            "final_state_readback": state_readback_s,
            "final_acceleration_readback": accel_readback_s,
            "direct_probe_reference": direct_s,
            "gpu_oracles_total": gpu_oracle_s,
            "total_wall": time.perf_counter() - wall0,
        },
        "scope": "BH #2D resident step on gfx950: f64 root bounds and Morton keys, onesweep 8-bit radix "
                 "sort, search-free sparse topology with preorder cells and a node stream, bottom-up f64 "
                 "aggregates, stackless node-stream traversal with correctly rounded f32 arithmetic, "
                 "kick/drift and final kick, all on one HIP stream with no host readback inside the loop.",
    }
    tree.close()
    evolving.close()
    runtime.close()
    gpu.close()
    return receipt


def main(argv=None):
    a = parse_args(argv)
    try:
        receipt = run(a)
    except BhError as e:
        print(f"GALAXY BH #2D: {e}", file=sys.stderr)
        return 1
    text = json.dumps(receipt, indent=2)
    if a.receipt:
        d = os.path.dirname(a.receipt)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(a.receipt, "w") as f:
            f.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
