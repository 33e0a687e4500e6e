"""Fail-closed validation of BH #2D receipts (galaxy.barnes-hut-parallel-gpu-tree-receipt.v1).

Restates the invariants of validate_receipt (scripts/bench-bh2d-hardware.py:1013-1390)
for receipts of galaxy_amd.driver: workload echo, hardware class, zero host
rebuilds/readbacks, steps + 1 force solves and builds, frozen ordering and layout
declarations, checksum formats and the repeat-rebuild match, the quadtree shape
bounds on cell, leaf and depth counts, the direct-probe gates and RMS/max
consistency, the benchmark sample count and upper median, and the frozen
sparse-layout byte formula.

Differences from the reference validator (this build's receipts record them):
  * backend "ROCm-HIP" is accepted beside Vulkan/Metal/Dx12;
  * the CPU BH #2A f64 sections (final_force.bh2a_flat_status and
    trajectory_vs_bh2a_flat_f64) may carry "skipped-test-oracle" (the CPU
    restatement is test infrastructure; tests/ attach its results to a receipt
    and validate the executed form with the reference gates).
The GPU oracles and the BH #2C benchmark follow the reference's rules.
"""
from __future__ import annotations

import json
import math
import re
from typing import Any

RECEIPT_SCHEMA = "galaxy.barnes-hut-parallel-gpu-tree-receipt.v1"
BH2C_CAP = 4096
TREE_LEVELS = 17
TREE_BUCKET_SIZE = 4
# Frozen labels: the validator of scripts/bench-bh2d-hardware.py:40-55 demands these
# strings verbatim.  They name the reference's tree (4-bit LSD radix, bind groups),
# not the kernels that produce this receipt (8-bit onesweep passes, HIP graphs); the
# ordering and layout they define are what the receipt's evidence follows.
TREE_ORDERING ="stable-lsd-radix-4bit-morton-code; resident body index retained for equal Morton keys"
TREE_LAYOUT = "sparse-level-order-slot=depth*N+group_start"
BENCHMARK_SAMPLE_SCOPE = ("complete rebuild call including host-side uniform/bind-group setup, "
                          "command encoding, submit and synchronization")
BACKENDS = {"Vulkan", "Metal", "Dx12", "ROCm-HIP"}
TEST_ORACLE_SKIP = "skipped-test-oracle"


class ReceiptError(ValueError):
    pass


def require(cond: bool, msg: str):
    if not cond:
        raise ReceiptError(msg)


def require_object(v: Any, name: str) -> dict:
    require(isinstance(v, dict), f"{name} must be an object")
    return v


def require_int(v: Any, name: str, minimum: int | None = None) -> int:
    require(isinstance(v, int) and not isinstance(v, bool), f"{name} must be an integer")
    if minimum is not None:
        require(v >= minimum, f"{name} must be >= {minimum}")
    return v


def require_number(v: Any, name: str, nonnegative=False, positive=False) -> float:
    require(isinstance(v, (int, float)) and not isinstance(v, bool) and math.isfinite(v),
            f"{name} must be a finite number")
    if nonnegative:
        require(v >= 0, f"{name} must be nonnegative")
    if positive:
        require(v > 0, f"{name} must be positive")
    return float(v)


def require_string(v: Any, name: str, nonempty=False) -> str:
    require(isinstance(v, str), f"{name} must be a string")
    if nonempty:
        require(bool(v.strip()), f"{name} must be nonempty")
    return v


def require_checksum(v: Any, name: str) -> str:
    s = require_string(v, name)
    require(re.fullmatch(r"[0-9a-f]{16}", s) is not None, f"{name} must be 16 lowercase hex digits")
    return s


def require_same_number(v: Any, expected: float, name: str) -> float:
    x = require_number(v, name)
    require(x == float(expected), f"{name} does not match requested workload")
    return x


def upper_median(values):
    v = sorted(values)
    return v[len(v) // 2]


def validated_samples(v: Any, name: str, count: int) -> list:
    require(isinstance(v, list), f"{name} must be a list")
    require(len(v) == count, f"{name} sample count mismatch")
    return [require_number(x, f"{name}[{i}]", positive=True) for i, x in enumerate(v)]


def expected_parallel_tree_buffer_bytes(particles: int) -> int:
    """The reference's sparse-layout allocation (nbody_parallel.rs:485-491, 990-992)."""
    blocks = (particles + 127) // 128
    return (particles * 32 + 2 * particles * 16 + particles * TREE_LEVELS * 64 + 32 + 2 * blocks * 16
            + 2 * blocks * 16 * 4)


def validate_state_error(v: Any, name: str) -> dict:
    s = require_object(v, name)
    keys = ("position_rms_relative_l2", "position_max_relative", "velocity_rms_relative_l2",
            "velocity_max_relative")
    r = {k: require_number(s.get(k), f"{name}.{k}", nonnegative=True) for k in keys}
    require((r[keys[0]] == 0.0) == (r[keys[1]] == 0.0), f"{name} position RMS/max zero-state mismatch")
    require((r[keys[2]] == 0.0) == (r[keys[3]] == 0.0), f"{name} velocity RMS/max zero-state mismatch")
    require(r[keys[0]] < 0.03 and r[keys[2]] < 0.03, f"{name} RMS gate failed")
    require(r[keys[1]] < 0.30 and r[keys[3]] < 0.30, f"{name} max gate failed")
    return r


def validate_force_errors(v: Any, name: str, samples: int):
    e = require_object(v, name)
    rms = require_number(e.get("force_rms_relative"), f"{name}.force_rms_relative", nonnegative=True)
    mx = require_number(e.get("force_max_relative"), f"{name}.force_max_relative", nonnegative=True)
    require(rms <= mx, f"{name} force RMS cannot exceed force maximum")
    require(mx <= rms * math.sqrt(samples), f"{name} force maximum is too large for RMS and sample count")
    require(rms < 0.04 and mx < 0.30, f"{name} force gate failed")


def adapter_identity(gpu: Any) -> str:
    info = require_object(gpu, "receipt.gpu")
    index = require_int(info.get("index"), "receipt.gpu.index", minimum=0)
    count = require_int(info.get("adapter_count"), "receipt.gpu.adapter_count", minimum=1)
    require(index < count, "receipt.gpu.index must be below adapter_count")
    name = require_string(info.get("name"), "receipt.gpu.name", nonempty=True)
    device_type = require_string(info.get("device_type"), "receipt.gpu.device_type", nonempty=True)
    soft = device_type.lower() == "cpu" or any(m in name.lower() for m in
                                               ("llvmpipe", "lavapipe", "swiftshader", "software"))
    require(not soft, "receipt.gpu contradicts the producer software-adapter classification")
    backend = require_string(info.get("backend"), "receipt.gpu.backend", nonempty=True)
    require(backend in BACKENDS, "receipt.gpu.backend is not enabled by the producer")
    driver = require_string(info.get("driver"), "receipt.gpu.driver")
    driver_info = require_string(info.get("driver_info"), "receipt.gpu.driver_info")
    require(bool(driver.strip() or driver_info.strip()), "receipt.gpu must include driver provenance")
    require(info.get("software") is False, "receipt.gpu.software must be false for hardware evidence")
    return json.dumps({"index": index, "adapter_count": count, "name": name, "backend": backend,
                       "device_type": device_type, "driver": driver, "driver_info": driver_info},
                      sort_keys=True, separators=(",", ":"))


def validate_tree_shape(tree: dict, particles: int) -> tuple:
    active = require_int(tree.get("active_cell_count"), "receipt.tree.active_cell_count", minimum=1)
    leaves = require_int(tree.get("leaf_count"), "receipt.tree.leaf_count", minimum=1)
    maxd = require_int(tree.get("max_depth"), "receipt.tree.max_depth", minimum=0)
    require(active <= TREE_LEVELS * particles, "active_cell_count exceeds the frozen sparse-cell capacity")
    require(leaves <= active, "leaf_count cannot exceed active_cell_count")
    internal = active - leaves
    require(leaves <= 3 * internal + 1, "leaf_count exceeds quadtree fan-out capacity")
    require(leaves <= particles, "leaf_count cannot exceed the resident particle count")
    require(maxd <= TREE_LEVELS - 1, "max_depth exceeds the frozen Morton depth")
    per_depth = particles // (TREE_BUCKET_SIZE + 1)
    require(active <= leaves + sum(min(4 ** d, per_depth) for d in range(maxd)),
            "active_cell_count exceeds per-depth quadtree occupancy")
    require(active >= maxd + 1, "active_cell_count is too small for the reported maximum depth")
    require(internal >= maxd, "too few internal cells for the reported maximum depth")
    require(internal <= sum(min(4 ** d, leaves) for d in range(maxd)),
            "too many internal cells for the leaf-path capacity")
    if maxd < TREE_LEVELS - 1:
        require(leaves >= (particles + TREE_BUCKET_SIZE - 1) // TREE_BUCKET_SIZE,
                "leaf_count is too small for the frozen bucket size below maximum depth")
    if particles > TREE_BUCKET_SIZE:
        require(active > leaves, "the tree must contain an internal cell above the bucket size")
        require(maxd > 0, "max_depth must be positive above the bucket size")
    return active, leaves, maxd


def validate_receipt(receipt: Any, *, particles: int, preset: str, steps: int, dt_myr: float, seed: int,
                     theta: float, softening_kpc: float, direct_probes: int, oracle_limit: int,
                     benchmark_warmup: int, benchmark_repeats: int) -> dict:
    root = require_object(receipt, "receipt")
    require(root.get("schema") == RECEIPT_SCHEMA, "unexpected BH #2D receipt schema")
    require(root.get("status") == "complete", "BH #2D receipt is not complete")
    require(root.get("phase") == "BH-2D", "receipt phase is not BH-2D")
    require(root.get("builder") == "gpu-parallel-sparse-radix-v1", "unexpected BH #2D builder")
    require(root.get("preset") == preset, "receipt preset does not match requested workload")
    require(require_int(root.get("particles"), "receipt.particles") == particles,
            "receipt particle count does not match")
    require(require_int(root.get("steps"), "receipt.steps") == steps, "receipt step count does not match")
    require_same_number(root.get("dt_myr"), dt_myr, "receipt.dt_myr")
    require(require_int(root.get("seed"), "receipt.seed") == seed, "receipt seed does not match")
    require_same_number(root.get("theta"), theta, "receipt.theta")
    require_same_number(root.get("softening_kpc"), softening_kpc, "receipt.softening_kpc")
    require_same_number(root.get("simulated_time_myr"), dt_myr * steps, "receipt.simulated_time_myr")
    require(root.get("measurement_class") == "hardware", "software-validation receipt rejected")
    require(root.get("hardware_performance_claim_allowed") is True,
            "receipt does not authorize hardware performance evidence")
    require(require_int(root.get("host_tree_rebuilds"), "receipt.host_tree_rebuilds", 0) == 0,
            "host tree rebuild occurred")
    require(require_int(root.get("host_particle_readbacks_during_steps"),
                        "receipt.host_particle_readbacks_during_steps", 0) == 0,
            "host particle readback occurred inside the evolution loop")
    require(require_int(root.get("force_solves"), "receipt.force_solves", 0) == steps + 1,
            "unexpected force-solve count")
    require(require_int(root.get("evolution_tree_builds"), "receipt.evolution_tree_builds", 0) == steps + 1,
            "unexpected tree-build count")
    identity = adapter_identity(root.get("gpu"))

    tree = require_object(root.get("tree"), "receipt.tree")
    require(tree.get("ordering") == TREE_ORDERING, "receipt.tree.ordering does not match the frozen ordering")
    require(tree.get("layout") == TREE_LAYOUT, "receipt.tree.layout does not match the frozen layout")
    require(tree.get("repeat_rebuild_matches") is True, "same-state repeat tree checksum changed")
    require_checksum(tree.get("initial_checksum_fnv_mix64"), "receipt.tree.initial_checksum_fnv_mix64")
    final = require_checksum(tree.get("final_checksum_fnv_mix64"), "receipt.tree.final_checksum_fnv_mix64")
    repeat = require_checksum(tree.get("repeat_checksum_fnv_mix64"), "receipt.tree.repeat_checksum_fnv_mix64")
    require(final == repeat, "receipt repeat tree checksum does not match the final tree checksum")
    active, leaves, maxd = validate_tree_shape(tree, particles)

    force = require_object(root.get("final_force"), "receipt.final_force")
    probes = min(direct_probes, particles)
    require(require_int(force.get("direct_probe_count"), "receipt.final_force.direct_probe_count") == probes,
            "receipt direct-probe count does not match requested workload")
    rms = require_number(force.get("direct_probe_rms_relative"), "direct_probe_rms_relative", nonnegative=True)
    mx = require_number(force.get("direct_probe_max_relative"), "direct_probe_max_relative", nonnegative=True)
    require(rms <= mx, "direct-force RMS cannot exceed maximum")
    require(mx <= rms * math.sqrt(probes), "direct-force maximum is too large for RMS and probe count")
    require(rms < 0.04, "direct-force RMS gate failed")
    require(mx < 0.30, "direct-force max gate failed")

    expected = "executed" if particles <= oracle_limit else "skipped-particle-limit"
    status = force.get("bh2a_flat_status")
    require(status in (expected, TEST_ORACLE_SKIP), f"BH #2A final-force status mismatch: expected {expected}")
    if status == "executed":
        frms = require_number(force.get("gpu_vs_bh2a_flat_rms_relative"), "gpu_vs_bh2a_flat_rms_relative",
                              nonnegative=True)
        fmx = require_number(force.get("gpu_vs_bh2a_flat_max_relative"), "gpu_vs_bh2a_flat_max_relative",
                             nonnegative=True)
        require(frms <= fmx, "BH #2A final-force RMS cannot exceed maximum")
        require(fmx <= frms * math.sqrt(particles), "BH #2A final-force maximum too large for RMS")
        require(frms < 0.04 and fmx < 0.30, "BH #2A final-force gate failed")
    oracles = require_object(root.get("gpu_oracles"), "receipt.gpu_oracles")
    require(oracles.get("status") == expected, f"GPU oracle status mismatch: expected {expected}")
    if oracles.get("status") == "executed":
        for key in ("bh2c_serial_gpu", "bh2b2_host_tree_gpu"):
            o = require_object(oracles.get(key), f"receipt.gpu_oracles.{key}")
            validate_state_error(o.get("state_error"), f"receipt.gpu_oracles.{key}.state_error")
            validate_force_errors(o, f"receipt.gpu_oracles.{key}", particles)
    else:
        require(require_int(oracles.get("limit"), "receipt.gpu_oracles.limit") == oracle_limit,
                "GPU oracle skip limit does not match requested oracle limit")
    traj = require_object(root.get("trajectory_vs_bh2a_flat_f64"), "receipt.trajectory_vs_bh2a_flat_f64")
    require(traj.get("status") in (expected, TEST_ORACLE_SKIP), "CPU trajectory status mismatch")
    if traj.get("status") == "executed":
        validate_state_error(traj.get("state_error"), "receipt.trajectory_vs_bh2a_flat_f64.state_error")
    else:
        require(require_int(traj.get("limit"), "receipt.trajectory_vs_bh2a_flat_f64.limit") == oracle_limit,
                "CPU trajectory skip limit does not match requested oracle limit")

    bench = require_object(root.get("tree_build_benchmark"), "receipt.tree_build_benchmark")
    require(require_string(bench.get("sample_scope"), "sample_scope", nonempty=True) == BENCHMARK_SAMPLE_SCOPE,
            "benchmark sample scope does not match the frozen complete-rebuild declaration")
    require(bench.get("stage_timings_are_diagnostics") is True, "benchmark must mark stage timings as diagnostics")
    require(require_int(bench.get("warmup"), "warmup") == benchmark_warmup, "benchmark warmup count mismatch")
    require(require_int(bench.get("repeats"), "repeats") == benchmark_repeats, "benchmark repeat count mismatch")
    samples = validated_samples(bench.get("parallel_samples_seconds"), "parallel_samples_seconds",
                                benchmark_repeats)
    median = upper_median(samples)
    require(require_number(bench.get("parallel_median_seconds"), "parallel_median_seconds", positive=True)
            == median, "parallel benchmark median does not match its samples")
    serial = require_object(bench.get("bh2c_serial"), "receipt.tree_build_benchmark.bh2c_serial")
    speedup = None
    if serial.get("status") == "executed":
        require(particles <= BH2C_CAP, "BH #2C comparison must skip above 4096")
        s = validated_samples(serial.get("samples_seconds"), "bh2c_serial.samples_seconds", benchmark_repeats)
        smed = upper_median(s)
        require(require_number(serial.get("median_seconds"), "bh2c_serial.median_seconds", positive=True) == smed,
                "BH #2C benchmark median does not match its samples")
        speedup = smed / median
        require(require_number(serial.get("parallel_vs_serial_speedup"), "parallel_vs_serial_speedup",
                               positive=True) == speedup, "BH #2C speedup does not match the medians")
    elif particles > BH2C_CAP:
        require(serial.get("status") == "skipped-bh2c-cap", "BH #2C comparison must skip above 4096")
        require(require_int(serial.get("limit"), "receipt.tree_build_benchmark.bh2c_serial.limit") == BH2C_CAP,
                "BH #2C benchmark skip limit is invalid")
    else:
        require(False, "BH #2C comparison should execute at this size")

    tb = require_int(root.get("parallel_tree_buffer_bytes"), "receipt.parallel_tree_buffer_bytes", minimum=1)
    require(tb == expected_parallel_tree_buffer_bytes(particles),
            "receipt.parallel_tree_buffer_bytes does not match the frozen BH #2D allocation formula")
    return {"particles": particles, "parallel_tree_buffer_bytes": tb, "active_cell_count": active,
            "leaf_count": leaves, "max_depth": maxd, "parallel_median_seconds": median,
            "bh2c_parallel_vs_serial_speedup": speedup, "direct_probe_rms_relative": rms,
            "direct_probe_max_relative": mx, "adapter_identity": identity}
