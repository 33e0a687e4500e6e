"""The BH #2D driver (galaxy_amd.driver, galaxy-bh-gpu-tree-parallel semantics) on
the GPU: its receipt passes the fail-closed validation, its tree checksums equal
the oracle's on the oracle's own trajectory, and with the oracle's BH #2A force
and trajectory comparisons attached the receipt passes the reference gates."""
import numpy as np
import pytest

from galaxy_amd import driver
from galaxy_amd import receipt as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset,n,steps", [("disc", 512, 3), ("collision", 1500, 2)])
def test_driver_receipt(oracle, preset, n, steps):
    print(f"driver {preset} n={n} steps={steps}", flush=True)
    a = driver.parse_args(["--preset", preset, "--particles", str(n), "--steps", str(steps)])
    keep = {}
    r = driver.run(a, keep)
    workload = dict(particles=n, preset=preset, steps=steps, dt_myr=a.dt_myr, seed=a.seed, theta=a.theta,
                    softening_kpc=a.softening_kpc, direct_probes=a.direct_probes, oracle_limit=a.oracle_limit,
                    benchmark_warmup=a.benchmark_warmup, benchmark_repeats=a.benchmark_repeats)
    R.validate_receipt(r, **workload)

    cfg = oracle.make_config(theta=a.theta, softening=a.softening_kpc, g=driver.G_KPC3_PER_MSUN_MYR2,
                             bucket=4, max_depth=48)
    # the GPU trajectory is the oracle's f32 trajectory bit for bit, and so are the trees
    rs, ra = oracle.f32_evolve(keep["initial"], steps, a.dt_myr, cfg)
    assert np.array_equal(keep["final_state"].view(np.uint32), rs.view(np.uint32))
    assert np.array_equal(keep["final_accel"].view(np.uint32), ra.view(np.uint32))
    t0 = oracle.tree_from_state(keep["initial"], cfg)
    t1 = oracle.tree_from_state(rs, cfg)
    assert r["tree"]["initial_checksum_fnv_mix64"] == f"{oracle.parallel_tree_checksum(*t0.sparse_evidence()):016x}"
    assert r["tree"]["final_checksum_fnv_mix64"] == f"{oracle.parallel_tree_checksum(*t1.sparse_evidence()):016x}"

    # the executed oracle sections (galaxy-bh-gpu-tree-parallel.rs:387-420) with the reference gates
    s = rs.astype(np.float64)
    fx, fy, _ = t1.accelerations()
    fe = oracle.force_error(ra[:, 0], ra[:, 1], fx, fy)
    r["final_force"].update(bh2a_flat_status="executed", gpu_vs_bh2a_flat_rms_relative=fe["rms_relative"],
                            gpu_vs_bh2a_flat_max_relative=fe["max_relative"])
    i = keep["initial"].astype(np.float64)
    cpu = oracle.flat_evolve(i[:, 0], i[:, 1], i[:, 4], i[:, 5], i[:, 2], steps, a.dt_myr, cfg)
    e = oracle.state_error((s[:, 0], s[:, 1], s[:, 4], s[:, 5]), cpu)
    r["trajectory_vs_bh2a_flat_f64"] = {"status": "executed", "state_error": {
        "position_rms_relative_l2": e["position_rms"], "position_max_relative": e["position_max"],
        "velocity_rms_relative_l2": e["velocity_rms"], "velocity_max_relative": e["velocity_max"]}}
    R.validate_receipt(r, **workload)
