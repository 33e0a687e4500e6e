"""CPU pins of the oracle's composed functions: the two checksums and the two
trajectories.

Each is checked against an independent restatement here, built from parts that
other tests pin (the flat tree, its f64 traversal, the packing and the f32
traversal) or from the definition itself:

- state_checksum (nbody/src/lib.rs:530-543) and flat_topology_checksum
  (nbody/src/flat.rs:410-433): a Python FNV-mix64 over the word sequence each
  defines;
- cpu_flat_evolve (runtime/src/bin/galaxy-bh-gpu-tree-parallel.rs:124-148): two
  steps composed in numpy from FlatTree.accelerations and the f64 KDK of that
  function;
- the drop-in f32 trajectory: two steps composed from tree_from_state, pack,
  traverse_f32 and the f32 KDK of runtime/src/nbody_evolve.wgsl:19-46 in numpy.

Each test also plants a plausible mistake in the restatement (a dropped or
reordered word, a missing half kick, a swapped operand) and checks that it no
longer matches, so the comparison would catch that mistake in the oracle.
"""
import numpy as np
import pytest

from galaxy_amd import ics

NO_CHILD = 0xFFFFFFFF
MASK = (1 << 64) - 1
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x00000100000001B3


def mix(h, w):
    """checksum_word (nbody/src/flat.rs:410-414): h ^= w; h *= prime; h ^= rotl(w, 23)."""
    h ^= w
    h = (h * FNV_PRIME) & MASK
    h ^= ((w << 23) | (w >> 41)) & MASK
    return h


def fold(words, h=FNV_OFFSET):
    for w in words:
        h = mix(h, int(w))
    return h


def f64_bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


# ------------------------------------------------------------- state_checksum
def state_words(x, y, vx, vy, m):
    """lib.rs:533-538: per body x, y, vx, vy, mass as f64 bit patterns."""
    cols = [f64_bits(v) for v in (x, y, vx, vy, m)]
    return np.stack(cols, axis=1).reshape(-1)


def test_state_checksum_matches_definition(oracle):
    b = ics.reference_collision(97, 41)
    x, y, vx, vy, m = b.T
    assert oracle.state_checksum(x, y, vx, vy, m) == fold(state_words(x, y, vx, vy, m))
    # edge words: signed zeros, subnormals, huge values, NaN payload
    e = np.array([[0.0, -0.0, 5e-324, 1.7976931348623157e308, np.nan],
                  [-1.5, 2.25, -0.0, 1e-300, 3.0]])
    assert oracle.state_checksum(*e.T) == fold(state_words(*e.T))
    assert oracle.state_checksum(*np.zeros((0, 5)).T) == FNV_OFFSET
    # planted mistakes: velocities swapped, one word dropped
    assert oracle.state_checksum(x, y, vx, vy, m) != fold(state_words(x, y, vy, vx, m))
    assert oracle.state_checksum(x, y, vx, vy, m) != fold(state_words(x, y, vx, vy, m)[:-1])


def test_fnv_mix_known_value():
    """One word from the offset, worked by hand: 1 ^ offset, times the prime
    (mod 2^64), xor 1 << 23."""
    h = FNV_OFFSET ^ 1
    h = (h * FNV_PRIME) % (1 << 64)
    h ^= 1 << 23
    assert mix(FNV_OFFSET, 1) == h
    # rotate-left by 23 of a word with its top bit set wraps that bit to bit 22
    assert mix(0, 1 << 63) == ((((1 << 63) * FNV_PRIME) & MASK) ^ (1 << 22))


# ------------------------------------------------------ flat_topology_checksum
def topology_words(tree):
    """flat.rs:416-433: morton_axis_bits (16), body_count, every entry's (code,
    body_index), every cell's (start, end, depth, children[0..4])."""
    code, idx, _ = tree.entries()
    _, topo = tree.cells()
    words = [16, tree.n]
    words += np.stack([code, idx], axis=1).reshape(-1).tolist()
    words += topo[:, [0, 1, 2, 3, 4, 5, 6]].reshape(-1).tolist()
    return words


@pytest.mark.parametrize("n,bucket,max_depth", [(2, 1, 16), (300, 4, 48), (1200, 1, 3)])
def test_flat_topology_checksum_matches_definition(oracle, n, bucket, max_depth):
    b = ics.reference_disc(n, 11 + n, 5.0e10, 3.0)
    cfg = oracle.make_config(bucket=bucket, max_depth=max_depth)
    t = oracle.FlatTree(b[:, 0], b[:, 1], b[:, 4], cfg)
    words = topology_words(t)
    _, topo = t.cells()
    assert (topo[:, 3:7] == NO_CHILD).any()  # absent children hash as u32::MAX
    assert t.topology_checksum() == fold(words)
    # planted mistakes: axis bits left out, a cell's child order reversed
    assert t.topology_checksum() != fold(words[1:])
    if len(topo) > 1:
        bad = topo.copy()
        bad[0, 3:7] = bad[0, 3:7][::-1]
        code, idx, _ = t.entries()
        alt = [16, t.n] + np.stack([code, idx], 1).reshape(-1).tolist() + bad.reshape(-1).tolist()
        assert t.topology_checksum() != fold(alt)


# ------------------------------------------------------------ cpu_flat_evolve
def numpy_flat_evolve(oracle, x, y, vx, vy, m, steps, dt, cfg, half_kick=0.5):
    """galaxy-bh-gpu-tree-parallel.rs:124-148 composed from the pinned f64 flat
    traversal: force, then per step kick/drift, rebuild, force, final kick."""
    x, y, vx, vy = (np.array(v, np.float64) for v in (x, y, vx, vy))
    ax, ay, _ = oracle.FlatTree(x, y, m, cfg).accelerations()
    for _ in range(steps):
        vx = vx + half_kick * dt * ax
        vy = vy + half_kick * dt * ay
        x = x + dt * vx
        y = y + dt * vy
        ax, ay, _ = oracle.FlatTree(x, y, m, cfg).accelerations()
        vx = vx + half_kick * dt * ax
        vy = vy + half_kick * dt * ay
    return x, y, vx, vy


@pytest.mark.parametrize("preset,n,seed", [("disc", 512, 303), ("collision", 700, 9)])
def test_flat_evolve_two_steps_bit_exact(oracle, preset, n, seed):
    b = ics.reference_disc(n, seed, 5.0e10, 3.0) if preset == "disc" else ics.reference_collision(n, seed)
    x, y, vx, vy, m = b.T
    cfg = oracle.make_config()
    dt = 0.5
    got = oracle.flat_evolve(x, y, vx, vy, m, 2, dt, cfg)
    want = numpy_flat_evolve(oracle, x, y, vx, vy, m, 2, dt, cfg)
    for g, w in zip(got, want):
        assert np.array_equal(f64_bits(g), f64_bits(w))
    # the state moved, and a dropped half kick would have been seen
    assert not np.array_equal(got[0], x)
    wrong = numpy_flat_evolve(oracle, x, y, vx, vy, m, 2, dt, cfg, half_kick=1.0)
    assert not np.array_equal(f64_bits(got[2]), f64_bits(wrong[2]))
    # zero steps return the input unchanged
    same = oracle.flat_evolve(x, y, vx, vy, m, 0, dt, cfg)
    for g, w in zip(same, (x, y, vx, vy)):
        assert np.array_equal(f64_bits(g), f64_bits(w))


# ---------------------------------------------------- drop-in f32 trajectory
def numpy_kick_drift(s, a, dt):
    """bh_kick_drift (nbody_evolve.wgsl:19-33) in f32: v' = v + (0.5 dt) a; x += dt v'."""
    dt = np.float32(dt)
    h = np.float32(0.5) * dt
    hv = s[:, 4:6] + h * a[:, 0:2]
    s[:, 0:2] = s[:, 0:2] + dt * hv
    s[:, 4:6] = hv


def numpy_final_kick(s, a, dt):
    """bh_final_kick (nbody_evolve.wgsl:35-46) in f32."""
    h = np.float32(0.5) * np.float32(dt)
    s[:, 4:6] = s[:, 4:6] + h * a[:, 0:2]


def composed_f32_evolve(oracle, state, steps, dt, cfg, swap_order=False):
    s = np.array(state, np.float32)

    def force(st):
        return oracle.traverse_f32(*oracle.tree_from_state(st, cfg).pack())

    a = force(s)
    for _ in range(steps):
        if swap_order:  # planted mistake: positions drift before the half kick
            s[:, 0:2] = s[:, 0:2] + np.float32(dt) * s[:, 4:6]
            s[:, 4:6] = s[:, 4:6] + np.float32(0.5) * np.float32(dt) * a[:, 0:2]
        else:
            numpy_kick_drift(s, a, dt)
        a = force(s)
        numpy_final_kick(s, a, dt)
    return s, a


@pytest.mark.parametrize("kind,n", [("disc", 768), ("collapse", 2000)])
def test_f32_evolve_two_steps_bit_exact(oracle, kind, n):
    if kind == "disc":
        s0 = ics.rotating_disc(n, 303, softening=0.01)
        cfg = oracle.make_config(theta=0.5, softening=0.01, g=1.0, bucket=1, max_depth=16)
        dt = 1e-3
    else:
        s0 = ics.uniform_collapse(n, 77)
        cfg = oracle.make_config(theta=0.5, softening=0.005, g=1.0, bucket=4, max_depth=16)
        dt = 5e-4
    s, a = oracle.f32_evolve(s0, 2, dt, cfg)
    rs, ra = composed_f32_evolve(oracle, s0, 2, dt, cfg)
    assert np.array_equal(s.view(np.uint32), rs.view(np.uint32))
    assert np.array_equal(a.view(np.uint32), ra.view(np.uint32))
    assert not np.array_equal(s[:, 0:2], s0[:, 0:2])
    ws, _ = composed_f32_evolve(oracle, s0, 2, dt, cfg, swap_order=True)
    assert not np.array_equal(s.view(np.uint32), ws.view(np.uint32))
