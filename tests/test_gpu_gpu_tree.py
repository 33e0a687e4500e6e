"""BH #2C correctness-first GPU tree (gbh_gpu_tree_*) against the oracle's
restatement of nbody_tree.wgsl: the tree evidence (meta, sorted entries,
level-order cells), the forces of bh_traverse over that tree and whole
trajectories, bit for bit; plus the reference's errors."""
import numpy as np
import pytest

from galaxy_amd import BhError, Config, NbodyGpu, ics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    g = NbodyGpu()
    yield g
    g.close()


def disc(n, seed):
    return ics.rotating_disc(n, seed, softening=0.01)


def cases():
    rng = np.random.default_rng(11)
    out = {}
    for n in (2, 3, 37, 300, 1000, 4096):
        out[f"disc{n}"] = disc(n, 100 + n)
    s = disc(500, 7)
    s[:60, 0:2] = s[0, 0:2]  # coincident bodies
    out["coincident500"] = s
    s = disc(64, 8)
    s[:, 0:2] = s[0, 0:2]  # every body at one point: the span floor
    out["allsame64"] = s
    s = disc(2000, 9)
    s[:, 0:2] *= np.float32(1e-3)
    s[:5, 0:2] += rng.normal(size=(5, 2)).astype(np.float32)  # clustered with outliers
    out["clustered2000"] = s
    return out


CASES = cases()


def check_tree(gpu, oracle, state, bucket, max_depth, theta=0.5, eps=0.05, g=1.0):
    ev = gpu.create_evolving_state(state)
    tree = gpu.create_gpu_tree(ev)
    cfg = Config(theta=theta, softening_kpc=eps, g=g, bucket=bucket, max_depth=max_depth)
    gpu.rebuild_gpu_tree(ev, tree, cfg)
    meta, entries, cells = gpu.read_gpu_tree_evidence(tree)
    want = oracle.GpuTreeBuild(state, bucket, max_depth)
    assert np.array_equal(meta, want.meta)
    assert np.array_equal(entries, want.entries)
    assert np.array_equal(cells, want.cells)
    gpu.force_from_gpu_tree(ev, tree, cfg)
    acc = gpu.read_evolving_accelerations(ev)
    ref = want.accelerations(theta, eps, g)
    assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))
    return gpu, ev, tree


@pytest.mark.parametrize("name", sorted(CASES))
def test_tree_and_forces_bit_exact(gpu, oracle, name):
    print(f"bh2c {name}", flush=True)
    check_tree(gpu, oracle, CASES[name], 4, 48)


@pytest.mark.parametrize("bucket,max_depth,theta,eps", [(1, 48, 0.5, 0.05), (8, 48, 0.7, 0.01), (4, 3, 0.5, 0.05),
                                                        (4, 16, 0.0, 0.05), (4, 48, 0.5, 0.0), (1024, 48, 0.5, 0.05)])
def test_config_variants(gpu, oracle, bucket, max_depth, theta, eps):
    print(f"bh2c bucket={bucket} max_depth={max_depth} theta={theta} eps={eps}", flush=True)
    check_tree(gpu, oracle, CASES["disc1000"], bucket, max_depth, theta, eps)


def test_theta_changed_between_rebuild_and_force(gpu, oracle):
    state = CASES["disc300"]
    ev = gpu.create_evolving_state(state)
    tree = gpu.create_gpu_tree(ev)
    gpu.rebuild_gpu_tree(ev, tree, Config(theta=0.5, softening_kpc=0.05, g=1.0, bucket=4, max_depth=48))
    gpu.force_from_gpu_tree(ev, tree, Config(theta=0.8, softening_kpc=0.02, g=2.0, bucket=4, max_depth=48))
    ref = oracle.GpuTreeBuild(state, 4, 48).accelerations(0.8, 0.02, 2.0)
    assert np.array_equal(gpu.read_evolving_accelerations(ev).view(np.uint32), ref.view(np.uint32))


def test_serial_oracle_trajectory_bit_exact(gpu, oracle):
    """run_serial_oracle (galaxy-bh-gpu-tree-parallel.rs:225-246) on the GPU
    equals the oracle's composition of the same steps."""
    state = CASES["disc1000"]
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=4, max_depth=48)
    dt, steps = 1e-3, 5
    print("bh2c trajectory", flush=True)
    ev = gpu.create_evolving_state(state)
    tree = gpu.create_gpu_tree(ev)
    gpu.rebuild_gpu_tree(ev, tree, cfg)
    gpu.force_from_gpu_tree(ev, tree, cfg)
    for _ in range(steps):
        gpu.kick_drift(ev, dt)
        gpu.rebuild_gpu_tree(ev, tree, cfg)
        gpu.force_from_gpu_tree(ev, tree, cfg)
        gpu.final_kick(ev, dt)
    s, a = oracle.gpu_tree_evolve(state, steps, dt, 0.5, 0.01, 1.0, 4, 48)
    assert np.array_equal(gpu.read_evolving_state(ev).view(np.uint32), s.view(np.uint32))
    assert np.array_equal(gpu.read_evolving_accelerations(ev).view(np.uint32), a.view(np.uint32))


def test_errors(gpu):
    big = gpu.create_evolving_state(disc(4097, 1))
    with pytest.raises(BhError, match=r"supports 2\.\.=4096 resident bodies"):
        gpu.create_gpu_tree(big)
    ev = gpu.create_evolving_state(CASES["disc37"])
    other = gpu.create_evolving_state(CASES["disc300"])
    tree = gpu.create_gpu_tree(ev)
    cfg = Config(theta=0.5, softening_kpc=0.05, g=1.0, bucket=4, max_depth=48)
    with pytest.raises(BhError, match="has not been built"):
        gpu.force_from_gpu_tree(ev, tree, cfg)
    with pytest.raises(BhError, match="count does not match"):
        gpu.rebuild_gpu_tree(other, tree, cfg)
    with pytest.raises(BhError, match="theta"):
        gpu.rebuild_gpu_tree(ev, tree, Config(theta=-1.0, softening_kpc=0.05, g=1.0, bucket=4, max_depth=48))
    with pytest.raises(BhError, match="cannot be represented as finite f32"):
        gpu.rebuild_gpu_tree(ev, tree, Config(theta=1e300, softening_kpc=0.05, g=1.0, bucket=4, max_depth=48))
    assert gpu.gpu_tree_buffer_bytes(tree) > 0
