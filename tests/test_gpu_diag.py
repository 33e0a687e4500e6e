"""GPU FP64 diagnostics against the oracle: f64 direct accelerations of probe
bodies (direct_acceleration_for, nbody/src/lib.rs:315-329), probe_error
(lib.rs:370-395) and the energy / Lz audit.  Both sides sum in the same order
with correctly rounded f64 square root and division, so the results are equal
bit for bit."""
import numpy as np
import pytest

from galaxy_amd import Config, NbodyGpu, ics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    g = NbodyGpu()
    yield g
    g.close()


def widened(state):
    s = state.astype(np.float64)
    return s[:, 0], s[:, 1], s[:, 2], s[:, 4], s[:, 5]


@pytest.mark.parametrize("case", ["disc768", "collapse3000", "clustered257"])
def test_direct_and_energy_match_oracle(gpu, oracle, case):
    if case == "disc768":
        state = ics.config0().states[0]
        cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=16)
    elif case == "collapse3000":
        state = ics.uniform_collapse(3000, 77)
        cfg = Config(theta=0.5, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)
    else:
        rng = np.random.default_rng(5)
        state = ics.uniform_collapse(257, 9)
        state[:40, 0:2] = state[0, 0:2]  # coincident bodies
        state[:, 4:6] = rng.normal(size=(257, 2)).astype(np.float32)
        cfg = Config(theta=0.5, softening_kpc=0.02, g=2.5, bucket=4, max_depth=16)
    print(f"diag {case} n={len(state)}", flush=True)
    ev = gpu.create_evolving_state(state)
    x, y, m, vx, vy = widened(state)
    n = len(state)
    idx = np.unique(np.r_[0, n - 1, np.arange(0, n, 7)]).astype(np.uint32)
    got = gpu.direct_f64(ev, cfg, idx)
    ocfg = oracle.make_config(theta=cfg.theta, softening=cfg.softening_kpc, g=cfg.g, bucket=cfg.bucket,
                              max_depth=cfg.max_depth)
    ex, ey = oracle.direct_for(x, y, m, ocfg, idx)
    assert np.array_equal(got[:, 0].view(np.uint64), np.asarray(ex).view(np.uint64))
    assert np.array_equal(got[:, 1].view(np.uint64), np.asarray(ey).view(np.uint64))
    e = gpu.energy(ev, cfg.softening_kpc, cfg.g)
    want = oracle.energy(x, y, vx, vy, m, cfg.softening_kpc, cfg.g)
    assert e == want
    # probe_error from the GPU direct sums equals the oracle's on the same accelerations
    rng = np.random.default_rng(1)
    approx = np.zeros((n, 2))
    full_idx = np.arange(n, dtype=np.uint32)
    fx, fy = oracle.direct_for(x, y, m, ocfg, full_idx)
    approx[:, 0] = np.asarray(fx) * (1.0 + 1e-3 * rng.normal(size=n))
    approx[:, 1] = np.asarray(fy) * (1.0 + 1e-3 * rng.normal(size=n))
    pe = gpu.probe_error(ev, cfg, approx, 64)
    want = oracle.probe_error(x, y, m, approx[:, 0], approx[:, 1], ocfg, 64)
    assert pe["rms_relative"] == pytest.approx(want["rms_relative"], rel=1e-12)
    assert pe["max_relative"] == pytest.approx(want["max_relative"], rel=1e-12)
