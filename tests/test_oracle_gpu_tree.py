"""Pins of the oracle's BH #2C tree (runtime/src/nbody_tree.wgsl:77-370).

Each part is checked against an independent restatement: numpy f32 bounds and
keys, a (code, index) lexicographic sort, a brute-force partition of the sorted
keys into prefix groups (level order = (depth, start) order), f32 halving for the
geometry, and f32 sums in the reference's order for the aggregates.  The
traversal over the tree is checked at theta = 0 against f64 direct summation and
at theta = 0.5 against the reference's BH #2D gates (nbody_gpu.rs tests use the
same bounds, galaxy-bh-gpu-tree-parallel.rs:398-403)."""
import numpy as np
import pytest

NO_CHILD = 0xFFFFFFFF


def state_of(x, y, m):
    s = np.zeros((len(x), 8), np.float32)
    s[:, 0], s[:, 1], s[:, 2] = x, y, m
    return s


def disc(n, seed):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(size=n))
    a = rng.uniform(0, 2 * np.pi, size=n)
    return state_of(r * np.cos(a), r * np.sin(a), np.full(n, 1.0 / n))


def f32(v):
    return np.float32(v)


def bounds_f32(s):
    x, y = s[:, 0], s[:, 1]
    span = max(max(f32(x.max() - x.min()), f32(y.max() - y.min())), f32(1.0e-12))
    cx = f32(f32(0.5) * f32(x.min() + x.max()))
    cy = f32(f32(0.5) * f32(y.min() + y.max()))
    half = f32(f32(f32(0.5) * span) + f32(1.0e-12))
    return cx, cy, half


def keys_f32(s, cx, cy, half):
    def q(v, c):
        low = np.float32(c - half)
        width = np.float32(np.float32(2.0) * half)
        scaled = np.floor(((v - low) / width) * np.float32(65536.0)).astype(np.float32)
        return np.clip(scaled, 0, 65535).astype(np.uint32)

    qx, qy = q(s[:, 0], cx), q(s[:, 1], cy)
    code = np.zeros(len(s), np.uint32)
    for b in range(16):
        code |= ((qx >> b) & 1) << (2 * b)
        code |= ((qy >> b) & 1) << (2 * b + 1)
    return code


def brute_cells(codes, bucket, max_depth):
    """Level-order cells (depth, start, end) of the sorted codes."""
    md = min(max_depth, 16)
    n = len(codes)
    out = [(0, 0, n)]
    level = [(0, 0, n)]
    while level:
        nxt = []
        for d, a, b in level:
            if b - a <= bucket or d >= md:
                continue
            shift = 30 - 2 * d
            k = a
            while k < b:
                qd = (int(codes[k]) >> shift) & 3
                e = k
                while e < b and ((int(codes[e]) >> shift) & 3) == qd:
                    e += 1
                nxt.append((d + 1, k, e))
                k = e
        out.extend(nxt)
        level = nxt
    return out


@pytest.mark.parametrize("n,bucket,max_depth,seed", [(2, 4, 48, 1), (37, 1, 48, 2), (300, 4, 48, 3),
                                                     (1000, 8, 5, 4), (4096, 4, 48, 5)])
def test_gpu_tree_matches_independent_restatement(oracle, n, bucket, max_depth, seed):
    s = disc(n, seed)
    t = oracle.GpuTreeBuild(s, bucket, max_depth)
    cx, cy, half = bounds_f32(s)
    assert t.bounds().tolist() == [cx, cy, half, 0.0]
    code = keys_f32(s, cx, cy, half)
    order = np.lexsort((np.arange(n), code))
    assert np.array_equal(t.entries[:, 0], code[order])
    assert np.array_equal(t.entries[:, 1], order.astype(np.uint32))
    assert np.array_equal(t.bodies[order, 4], np.arange(n, dtype=np.uint32))
    cells = brute_cells(code[order], bucket, max_depth)
    assert t.meta[0] == len(cells) and t.meta[3] == 0
    rd = t.cells[:, 8:11]
    assert [tuple(r) for r in rd[:, [2, 0, 1]].tolist()] == cells
    leaves = sum(1 for d, a, b in cells if b - a <= bucket or d >= min(max_depth, 16))
    assert t.meta[1] == leaves and t.meta[2] == max(d for d, _, _ in cells)
    # children: cell ids of the (depth + 1) groups inside the parent's range, by quadrant
    index = {(d, a): i for i, (d, a, b) in enumerate(cells)}
    geom = t.cells[:, 0:4].view(np.float32)
    com = t.cells[:, 4:8].view(np.float32)
    for i, (d, a, b) in enumerate(cells):
        ch = t.cells[i, 12:16]
        kids = [(q, c) for q, c in enumerate(ch.tolist()) if c != NO_CHILD]
        for q, c in kids:
            cd, ca, cb = cells[c]
            assert cd == d + 1 and a <= ca < cb <= b
            assert ((int(code[order][ca]) >> (30 - 2 * d)) & 3) == q
            h = f32(f32(0.5) * geom[i, 2])
            ox = h if q & 1 else -h
            oy = h if q & 2 else -h
            assert geom[c, 0] == f32(geom[i, 0] + ox) and geom[c, 1] == f32(geom[i, 1] + oy)
            assert geom[c, 2] == h
        assert sorted(index[(d + 1, cells[c][1])] for _, c in kids) == [c for _, c in kids]
    # aggregates in f32, in the reference order (reverse level order)
    mass = np.zeros(len(cells), np.float32)
    cxy = np.zeros((len(cells), 2), np.float32)
    for i in range(len(cells) - 1, -1, -1):
        d, a, b = cells[i]
        ch = [c for c in t.cells[i, 12:16].tolist() if c != NO_CHILD]
        m_ = wx = wy = f32(0.0)
        if not ch:
            for p in range(a, b):
                j = order[p]
                m_ = f32(m_ + s[j, 2])
                wx = f32(wx + f32(s[j, 2] * s[j, 0]))
                wy = f32(wy + f32(s[j, 2] * s[j, 1]))
        else:
            for c in ch:
                m_ = f32(m_ + mass[c])
                wx = f32(wx + f32(mass[c] * cxy[c, 0]))
                wy = f32(wy + f32(mass[c] * cxy[c, 1]))
        mass[i] = m_
        cxy[i] = (f32(wx / m_), f32(wy / m_))
    assert np.array_equal(geom[:, 3], mass)
    assert np.array_equal(com[:, 0:2], cxy)


def test_gpu_tree_coincident_bodies_chain_to_max_depth(oracle):
    s = state_of(np.full(9, 0.25), np.full(9, -0.5), np.full(9, 2.0))
    s[0, 0] = 1.0  # one body elsewhere so the bounds have a span
    t = oracle.GpuTreeBuild(s, 4, 48)
    deepest = t.cells[t.cells[:, 10] == t.meta[2]]
    assert t.meta[2] == 16
    assert len(deepest) == 1 and deepest[0, 9] - deepest[0, 8] == 8
    # coincident bodies keep their index order (the (code, index) order)
    a, b = deepest[0, 8], deepest[0, 9]
    assert t.entries[a:b, 1].tolist() == list(range(1, 9))


def test_gpu_tree_all_coincident_uses_the_span_floor(oracle):
    # span = max(0, 0, 1e-12): half = 0.5 * 1e-12 + 1e-12 in f32, every key the centre's
    s = state_of(np.full(6, 0.3), np.full(6, 0.3), np.full(6, 1.0))
    t = oracle.GpuTreeBuild(s, 4, 48)
    cx, cy, half = bounds_f32(s)
    assert half == f32(f32(f32(0.5) * f32(1.0e-12)) + f32(1.0e-12))
    assert t.bounds().tolist() == [cx, cy, half, 0.0]
    assert np.all(t.entries[:, 0] == keys_f32(s, cx, cy, half)[0])
    assert t.entries[:, 1].tolist() == list(range(6)) and t.meta[2] == 16


def test_gpu_tree_forces(oracle):
    s = disc(700, 8)
    t = oracle.GpuTreeBuild(s, 4, 48)
    x, y, m = (s[:, k].astype(np.float64) for k in range(3))
    cfg = oracle.make_config(theta=0.0, softening=0.05, g=1.0, bucket=4, max_depth=48)
    ex, ey = oracle.direct_accelerations(x, y, m, cfg)
    a0 = t.accelerations(0.0, 0.05, 1.0).astype(np.float64)
    e = oracle.force_error(a0[:, 0], a0[:, 1], ex, ey)
    assert e["max_relative"] < 1e-4
    a5 = t.accelerations(0.5, 0.05, 1.0).astype(np.float64)
    e = oracle.force_error(a5[:, 0], a5[:, 1], ex, ey)
    assert e["rms_relative"] < 0.04 and e["max_relative"] < 0.30
