"""Receipt validation (galaxy_amd.receipt) on synthetic receipts: a valid receipt
passes, and each fail-closed invariant of validate_receipt
(scripts/bench-bh2d-hardware.py:1013-1390) rejects a receipt that breaks it."""
import copy

import pytest

from galaxy_amd import receipt as R

WORKLOAD = dict(particles=512, preset="disc", steps=3, dt_myr=0.01, seed=303, theta=0.5,
                softening_kpc=0.05, direct_probes=12, oracle_limit=4096, benchmark_warmup=1,
                benchmark_repeats=5)


def receipt_for(n=512, steps=3):
    samples = [0.0011, 0.0009, 0.0012, 0.0010, 0.0013]
    serial = [0.004, 0.0041, 0.0039, 0.0042, 0.0040]
    return {
        "schema": R.RECEIPT_SCHEMA, "status": "complete", "phase": "BH-2D",
        "builder": "gpu-parallel-sparse-radix-v1", "measurement_class": "hardware",
        "hardware_performance_claim_allowed": True, "preset": "disc", "particles": n, "steps": steps,
        "dt_myr": 0.01, "simulated_time_myr": 0.01 * steps, "seed": 303, "theta": 0.5, "softening_kpc": 0.05,
        "gpu": {"index": 0, "adapter_count": 1, "name": "AMD Instinct MI355X", "backend": "ROCm-HIP",
                "device_type": "DiscreteGpu", "driver": "HIP 70226015", "driver_info": "gfx950",
                "software": False},
        "host_tree_rebuilds": 0, "host_particle_readbacks_during_steps": 0,
        "force_solves": steps + 1, "evolution_tree_builds": steps + 1,
        "parallel_tree_buffer_bytes": R.expected_parallel_tree_buffer_bytes(n),
        "tree": {"initial_checksum_fnv_mix64": "0123456789abcdef", "final_checksum_fnv_mix64": "fedcba9876543210",
                 "repeat_checksum_fnv_mix64": "fedcba9876543210", "repeat_rebuild_matches": True,
                 "active_cell_count": 280, "leaf_count": 200, "max_depth": 8,
                 "ordering": R.TREE_ORDERING, "layout": R.TREE_LAYOUT},
        "final_force": {"bh2a_flat_status": R.TEST_ORACLE_SKIP, "direct_probe_count": 12,
                        "direct_probe_rms_relative": 0.01, "direct_probe_max_relative": 0.02},
        "trajectory_vs_bh2a_flat_f64": {"status": R.TEST_ORACLE_SKIP, "limit": 4096},
        "gpu_oracles": {"status": "executed",
                        "bh2c_serial_gpu": {"state_error": small_state_error(), "force_rms_relative": 0.011,
                                            "force_max_relative": 0.05},
                        "bh2b2_host_tree_gpu": {"state_error": small_state_error(), "force_rms_relative": 0.002,
                                                "force_max_relative": 0.01}},
        "tree_build_benchmark": {"sample_scope": R.BENCHMARK_SAMPLE_SCOPE, "stage_timings_are_diagnostics": True,
                                 "warmup": 1, "repeats": 5, "parallel_samples_seconds": samples,
                                 "parallel_median_seconds": sorted(samples)[2],
                                 "bh2c_serial": {"status": "executed", "samples_seconds": serial,
                                                 "median_seconds": sorted(serial)[2],
                                                 "parallel_vs_serial_speedup": sorted(serial)[2] / sorted(samples)[2]}},
    }


def small_state_error():
    return {"position_rms_relative_l2": 1e-4, "position_max_relative": 1e-3,
            "velocity_rms_relative_l2": 2e-4, "velocity_max_relative": 3e-3}


def test_valid_receipt_passes():
    out = R.validate_receipt(receipt_for(), **WORKLOAD)
    assert out["active_cell_count"] == 280 and out["parallel_median_seconds"] == 0.0011


def test_executed_oracle_sections_are_gated():
    r = receipt_for()
    r["final_force"].update(bh2a_flat_status="executed", gpu_vs_bh2a_flat_rms_relative=0.01,
                            gpu_vs_bh2a_flat_max_relative=0.05)
    r["trajectory_vs_bh2a_flat_f64"] = {"status": "executed", "state_error": {
        "position_rms_relative_l2": 1e-6, "position_max_relative": 1e-5,
        "velocity_rms_relative_l2": 1e-6, "velocity_max_relative": 1e-5}}
    R.validate_receipt(r, **WORKLOAD)
    r["final_force"]["gpu_vs_bh2a_flat_max_relative"] = 0.31
    with pytest.raises(R.ReceiptError, match="final-force"):
        R.validate_receipt(r, **WORKLOAD)


def mutate(path, value):
    r = receipt_for()
    node = r
    for k in path[:-1]:
        node = node[k]
    if value is KeyError:
        del node[path[-1]]
    else:
        node[path[-1]] = value
    return r


@pytest.mark.parametrize("path,value,msg", [
    (("schema",), "other", "schema"),
    (("measurement_class",), "software-validation", "software-validation"),
    (("host_tree_rebuilds",), 1, "host tree rebuild"),
    (("host_particle_readbacks_during_steps",), 2, "readback"),
    (("force_solves",), 3, "force-solve"),
    (("evolution_tree_builds",), 5, "tree-build"),
    (("simulated_time_myr",), 0.02, "simulated_time_myr"),
    (("gpu", "software"), True, "software"),
    (("gpu", "backend"), "CUDA", "backend"),
    (("gpu", "name"), "llvmpipe", "software-adapter"),
    (("tree", "ordering"), "unstable", "ordering"),
    (("tree", "repeat_checksum_fnv_mix64"), "0000000000000000", "repeat tree checksum"),
    (("tree", "final_checksum_fnv_mix64"), "FEDCBA9876543210", "hex"),
    (("tree", "repeat_rebuild_matches"), False, "repeat"),
    (("tree", "leaf_count"), 281, "leaf_count"),
    (("tree", "leaf_count"), 100, "bucket size"),
    (("tree", "max_depth"), 17, "max_depth"),
    (("tree", "active_cell_count"), 200 + 8000, "capacity|occupancy"),
    (("final_force", "direct_probe_count"), 11, "direct-probe count"),
    (("final_force", "direct_probe_rms_relative"), 0.05, "RMS"),
    (("final_force", "direct_probe_max_relative"), 0.005, "RMS cannot exceed"),
    (("final_force", "direct_probe_max_relative"), 0.5, "maximum is too large|max gate"),
    (("final_force", "bh2a_flat_status"), "skipped-particle-limit", "final-force status"),
    (("gpu_oracles", "status"), R.TEST_ORACLE_SKIP, "GPU oracle status"),
    (("gpu_oracles", "bh2c_serial_gpu", "force_max_relative"), 0.31, "force gate|too large"),
    (("gpu_oracles", "bh2b2_host_tree_gpu", "state_error", "position_rms_relative_l2"), 0.05, "RMS gate"),
    (("gpu_oracles", "bh2c_serial_gpu", "force_rms_relative"), 0.06, "cannot exceed|force gate"),
    (("tree_build_benchmark", "bh2c_serial", "status"), "skipped-not-built", "should execute"),
    (("tree_build_benchmark", "bh2c_serial", "median_seconds"), 0.0041, "BH #2C benchmark median"),
    (("tree_build_benchmark", "bh2c_serial", "parallel_vs_serial_speedup"), 3.0, "speedup"),
    (("tree_build_benchmark", "parallel_median_seconds"), 0.0012, "median"),
    (("tree_build_benchmark", "parallel_samples_seconds"), [0.001], "sample count"),
    (("tree_build_benchmark", "stage_timings_are_diagnostics"), False, "diagnostics"),
    (("parallel_tree_buffer_bytes",), 12345, "allocation formula"),
])
def test_each_invariant_rejects(path, value, msg):
    with pytest.raises(R.ReceiptError, match=msg):
        R.validate_receipt(mutate(path, value), **WORKLOAD)


def test_sparse_layout_formula_matches_the_reference_example():
    # 17 * 65536 * 64 B of cells stays below 128 MiB (nbody_parallel.rs:1007-1012)
    assert 17 * 65536 * 64 < 128 << 20
    n = 65536
    blocks = n // 128
    assert R.expected_parallel_tree_buffer_bytes(n) == 32 * n + 32 * n + 1088 * n + 32 + 160 * blocks


def test_workload_echo_is_checked():
    r = receipt_for()
    bad = dict(WORKLOAD, seed=304)
    with pytest.raises(R.ReceiptError, match="seed"):
        R.validate_receipt(copy.deepcopy(r), **bad)


def test_library_checksum_is_fnv_mix64(oracle):
    """gbh_fnv_mix64 (host code of the library) against the closed form and the
    oracle's parallel_tree_checksum on the same evidence words."""
    import numpy as np

    from galaxy_amd import driver

    rng = np.random.default_rng(3)
    words = rng.integers(0, 1 << 32, size=777, dtype=np.uint64).astype(np.uint32)
    h = 0xCBF29CE484222325
    for w in words.tolist():
        h ^= w
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
        h ^= ((w << 23) | (w >> 41)) & 0xFFFFFFFFFFFFFFFF
    assert driver.checksum_words(words) == h
    meta = rng.integers(0, 1 << 32, size=8, dtype=np.uint64).astype(np.uint32)
    entries = rng.integers(0, 1 << 32, size=(33, 4), dtype=np.uint64).astype(np.uint32)
    cells = rng.integers(0, 1 << 32, size=(9, 16), dtype=np.uint64).astype(np.uint32)

    class Ev:
        pass

    ev = Ev()
    ev.meta, ev.entries, ev.cells = meta, entries, cells
    assert driver.parallel_tree_checksum(ev) == oracle.parallel_tree_checksum(meta, entries, cells)

def test_above_the_oracle_and_bh2c_caps():
    n = 5000
    r = receipt_for(n=n)
    r["tree"].update(active_cell_count=1800, leaf_count=1300, max_depth=9)
    r["parallel_tree_buffer_bytes"] = R.expected_parallel_tree_buffer_bytes(n)
    skip = {"status": "skipped-particle-limit", "limit": 4096}
    r["final_force"]["bh2a_flat_status"] = "skipped-particle-limit"
    r["trajectory_vs_bh2a_flat_f64"] = dict(skip)
    r["gpu_oracles"] = dict(skip)
    r["tree_build_benchmark"]["bh2c_serial"] = {"status": "skipped-bh2c-cap", "limit": 4096}
    w = dict(WORKLOAD, particles=n)
    assert R.validate_receipt(copy.deepcopy(r), **w)["bh2c_parallel_vs_serial_speedup"] is None
    r["tree_build_benchmark"]["bh2c_serial"]["limit"] = 1000
    with pytest.raises(R.ReceiptError, match="skip limit"):
        R.validate_receipt(copy.deepcopy(r), **w)
    r["tree_build_benchmark"]["bh2c_serial"] = {"status": "executed"}
    with pytest.raises(R.ReceiptError, match="skip above 4096"):
        R.validate_receipt(copy.deepcopy(r), **w)


def test_driver_metrics_equal_oracle(oracle):
    import numpy as np

    from galaxy_amd import driver

    rng = np.random.default_rng(9)
    a = rng.normal(size=(300, 8)).astype(np.float32)
    b = (a * (1 + 1e-3 * rng.normal(size=a.shape))).astype(np.float32)
    b[:5, 0:2] = 0.0  # positions at the origin: the 1e-6 floor
    got = driver.state_error(a, b)
    w = oracle.state_error((a[:, 0], a[:, 1], a[:, 4], a[:, 5]), (b[:, 0], b[:, 1], b[:, 4], b[:, 5]))
    for k, ok in (("position_rms_relative_l2", "position_rms"), ("position_max_relative", "position_max"),
                  ("velocity_rms_relative_l2", "velocity_rms"), ("velocity_max_relative", "velocity_max")):
        assert got[k] == pytest.approx(w[ok], rel=1e-12)
    fa, fb = a[:, :4], b[:, :4]
    rms, mx = driver.force_error(fa, fb)
    e = oracle.force_error(fa[:, 0], fa[:, 1], fb[:, 0], fb[:, 1])
    assert rms == pytest.approx(e["rms_relative"], rel=1e-12) and mx == e["max_relative"]
