"""CPU tests pinning the oracle (oracle/cpu_ref.c) before it is trusted.

Pins, in order: the reference tests' own known-answer data and accuracy gates
(nbody/src/lib.rs:545-635, nbody/src/flat.rs:435-598), exhaustive Morton
interleave, independent brute-force topology from the keys, closed-form
aggregates, an independent numpy f32 restatement of the theta=0 traversal sum,
and closed-form KDK.
"""
import json
import os

import numpy as np
import pytest

from galaxy_amd import ics

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
NO_CHILD = 0xFFFFFFFF


def rel_err(ax, ay, bx, by):
    return np.hypot(ax - bx, ay - by) / np.maximum(np.hypot(bx, by), 1e-30)


def disc(n, seed):
    b = ics.reference_disc(n, seed, 5.0e10, 3.0)
    return b[:, 0], b[:, 1], b[:, 4]


# ---------------------------------------------------------------- Morton keys
def test_morton_kats(oracle):
    lib = oracle.lib()
    for x, y, code in GOLD["morton2"]:
        assert lib.cref_morton2(x, y) == code
    o = GOLD["morton2_order"]
    assert (lib.cref_morton2(*o["a"]) < lib.cref_morton2(*o["b"])) == o["a_less_than_b"]


def test_morton_exhaustive_against_bit_loop(oracle):
    lib = oracle.lib()
    v = np.arange(65536, dtype=np.uint64)
    spread = np.zeros_like(v)
    for b in range(16):  # textbook bit-by-bit interleave
        spread |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
    got_x = np.array([lib.cref_morton2(int(i), 0) for i in range(65536)], dtype=np.uint64)
    got_y = np.array([lib.cref_morton2(0, int(i)) for i in range(65536)], dtype=np.uint64)
    assert np.array_equal(got_x, spread)
    assert np.array_equal(got_y, spread << np.uint64(1))
    rng = np.random.default_rng(1)
    for x, y in rng.integers(0, 65536, size=(2000, 2)):
        assert lib.cref_morton2(int(x), int(y)) == int(spread[x] | (spread[y] << np.uint64(1)))


def test_quantize_edges_and_numpy(oracle):
    lib = oracle.lib()
    c, h = 0.25, 3.5
    low, high = c - h, c + h
    assert lib.cref_quantize_axis(low, c, h) == 0
    assert lib.cref_quantize_axis(high, c, h) == 65535  # clamp of 65536
    assert lib.cref_quantize_axis(low - 10.0, c, h) == 0
    assert lib.cref_quantize_axis(high + 10.0, c, h) == 65535
    assert lib.cref_quantize_axis(c, c, h) == 32768
    vals = np.random.default_rng(2).uniform(low, high, 5000)
    expect = np.clip(np.floor((vals - (c - h)) / (h * 2.0) * 65536.0), 0, 65535).astype(np.int64)
    got = np.array([lib.cref_quantize_axis(float(v), c, h) for v in vals])
    assert np.array_equal(got, expect)


# ------------------------------------------------------ reference unit tests
def test_config_validation(oracle):
    bad = [dict(theta=-1.0), dict(theta=float("nan")), dict(softening=-0.1), dict(g=0.0),
           dict(bucket=0), dict(bucket=1025), dict(max_depth=0), dict(max_depth=129)]
    for kw in bad:
        cfg = oracle.make_config(**kw)
        assert oracle.lib().cref_config_validate(cfg) != 0, kw
    assert oracle.lib().cref_config_validate(oracle.make_config()) == 0


def test_tree_preserves_count_and_mass(oracle):
    x, y, m = disc(256, 303)
    t = oracle.FlatTree(x, y, m, oracle.make_config())
    geom, topo = t.cells()
    assert topo[0, 0] == 0 and topo[0, 1] == 256
    assert abs(geom[0, 3] - GOLD["gates"]["root_mass"]) < GOLD["gates"]["root_mass_tol"]
    c = t.counts()
    assert c["node_count"] > c["leaf_count"] and 0 < c["max_depth"] <= 16


def test_theta_zero_matches_direct(oracle):
    x, y, m = disc(128, 303)
    cfg = oracle.make_config(theta=0.0)
    dx, dy = oracle.direct_accelerations(x, y, m, cfg)
    ax, ay, _ = oracle.FlatTree(x, y, m, cfg).accelerations()
    assert rel_err(ax, ay, dx, dy).max() < GOLD["gates"]["theta0_vs_direct"]
    rx, ry, _ = oracle.recursive_accelerations(x, y, m, cfg)
    assert rel_err(rx, ry, dx, dy).max() < GOLD["gates"]["theta0_vs_direct"]


def test_theta_half_gate_and_flat_vs_recursive(oracle):
    x, y, m = disc(256, 303)
    cfg = oracle.make_config()
    dx, dy = oracle.direct_accelerations(x, y, m, cfg)
    ax, ay, st = oracle.FlatTree(x, y, m, cfg).accelerations()
    e = rel_err(ax, ay, dx, dy)
    assert np.sqrt(np.mean(e ** 2)) < GOLD["gates"]["theta_half_rms"]
    assert e.max() < GOLD["gates"]["theta_half_max"]
    assert st["approximated"] > 0
    rx, ry, _ = oracle.recursive_accelerations(x, y, m, cfg)
    assert rel_err(ax, ay, rx, ry).max() < GOLD["gates"]["flat_vs_recursive"]


def test_coincident_bodies_keep_resident_order(oracle):
    k = GOLD["coincident"]
    b = np.array([k["body"]] * k["count"])
    cfg = oracle.make_config(bucket=k["bucket"], max_depth=k["max_depth"], softening=k["softening"])
    t = oracle.FlatTree(b[:, 0], b[:, 1], b[:, 4], cfg)
    _, idx, _ = t.entries()
    assert idx.tolist() == k["order"]
    assert t.counts()["max_depth"] == k["tree_max_depth"]
    rx, ry, st = oracle.recursive_accelerations(b[:, 0], b[:, 1], b[:, 4],
                                                oracle.make_config(bucket=1, max_depth=12, softening=0.1))
    assert st["max_depth"] <= 12 and np.isfinite(rx).all() and np.isfinite(ry).all()


def test_max_depth_one_is_honoured(oracle):
    k = GOLD["max_depth_one"]
    b = np.zeros((k["count"], 5))
    b[:, 4] = 1.0
    cfg = oracle.make_config(bucket=k["bucket"], max_depth=k["max_depth"], softening=k["softening"])
    t = oracle.FlatTree(b[:, 0], b[:, 1], b[:, 4], cfg)
    _, topo = t.cells()
    assert t.counts()["max_depth"] == k["tree_max_depth"]
    assert (topo[:, 2] <= 1).all()


def test_membership_fixture(oracle):
    k = GOLD["membership"]
    b = np.array(k["bodies"])
    cfg = oracle.make_config(theta=k["theta"], bucket=k["bucket"], max_depth=k["max_depth"],
                             softening=k["softening"])
    ax, ay, _ = oracle.FlatTree(b[:, 0], b[:, 1], b[:, 4], cfg).accelerations()
    dx, dy = oracle.direct_accelerations(b[:, 0], b[:, 1], b[:, 4], cfg)
    assert rel_err(ax, ay, dx, dy).max() < k["max_relative_below"]


def test_checksums_repeatable(oracle):
    x, y, m = disc(512, 919)
    cfg = oracle.make_config()
    assert oracle.FlatTree(x, y, m, cfg).topology_checksum() == oracle.FlatTree(x, y, m, cfg).topology_checksum()
    a = ics.reference_collision(256, 999)
    b = ics.reference_collision(256, 999)
    assert np.array_equal(a, b)
    assert oracle.state_checksum(*a[:, [0, 1, 2, 3, 4]].T) == oracle.state_checksum(*b.T)


# ------------------------------------------------- independent brute force
def brute_topology(codes, bucket, maxd):
    """All (depth, start, end) cells from the sorted keys by prefix grouping."""
    n = len(codes)
    cells = []

    def rec(start, end, depth):
        cells.append((depth, start, end))
        if end - start <= bucket or depth >= maxd:
            return
        shift = 30 - 2 * depth
        q = (codes[start:end] >> shift) & 3
        bounds = np.flatnonzero(np.diff(q)) + 1
        edges = [start] + [start + int(b) for b in bounds] + [end]
        for a, b in zip(edges[:-1], edges[1:]):
            rec(a, b, depth + 1)

    rec(0, n, 0)
    return cells


@pytest.mark.parametrize("n,bucket,seed", [(300, 1, 5), (777, 4, 6), (2000, 8, 7), (64, 64, 8)])
def test_topology_matches_brute_force(oracle, n, bucket, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(0, 1, n // 2), rng.normal(3, 0.01, n - n // 2)])
    y = np.concatenate([rng.normal(0, 1, n // 2), rng.normal(-2, 0.01, n - n // 2)])
    m = rng.uniform(0.5, 2.0, n)
    cfg = oracle.make_config(bucket=bucket, max_depth=16)
    t = oracle.FlatTree(x, y, m, cfg)
    code, idx, pos = t.entries()
    # sort: a permutation, strictly ordered by (code, index), inverse positions
    assert sorted(idx.tolist()) == list(range(n))
    key = code.astype(np.uint64) << np.uint64(32) | idx.astype(np.uint64)
    assert (np.diff(key.astype(np.float64)) > 0).all() or (np.diff(key) > 0).all()
    assert np.array_equal(pos[idx], np.arange(n))
    geom, topo = t.cells()
    got = [(int(d), int(s), int(e)) for s, e, d in topo[:, :3]]
    assert got == brute_topology(code.astype(np.int64), bucket, 16)  # preorder
    # children partition the parent range in quadrant order; first child is k+1
    for k in range(len(topo)):
        ch = [c for c in topo[k, 3:] if c != NO_CHILD]
        if not ch:
            continue
        assert ch[0] == k + 1
        assert topo[ch[0], 0] == topo[k, 0] and topo[ch[-1], 1] == topo[k, 1]
        for a, b in zip(ch[:-1], ch[1:]):
            assert topo[a, 1] == topo[b, 0]
        for c in ch:
            assert topo[c, 2] == topo[k, 2] + 1
    # aggregates: mass and centre of mass of every cell from its own range
    xs, ys, ms = x[idx], y[idx], m[idx]
    for k in range(len(topo)):
        s, e = topo[k, 0], topo[k, 1]
        mm = ms[s:e].sum()
        assert abs(geom[k, 3] - mm) <= 1e-12 * mm
        assert abs(geom[k, 4] - (ms[s:e] * xs[s:e]).sum() / mm) <= 1e-9 * (1 + abs(geom[k, 4]))
        assert abs(geom[k, 5] - (ms[s:e] * ys[s:e]).sum() / mm) <= 1e-9 * (1 + abs(geom[k, 5]))
    # geometry: halving from the root along the key's quadrant path
    cx0, cy0, h0 = t.root()
    for k in range(len(topo)):
        d = topo[k, 2]
        c = int(code[topo[k, 0]])
        cx, cy, h = cx0, cy0, h0
        for level in range(d):
            q = (c >> (30 - 2 * level)) & 3
            h = h * 0.5
            cx = cx + (h if q & 1 else -h)
            cy = cy + (h if q & 2 else -h)
        assert (geom[k, 0], geom[k, 1], geom[k, 2]) == (cx, cy, h)


def test_root_bounds_closed_form(oracle):
    rng = np.random.default_rng(3)
    x, y = rng.normal(1, 2, 100), rng.normal(-1, 0.5, 100)
    t = oracle.FlatTree(x, y, np.ones(100), oracle.make_config())
    span = max(x.max() - x.min(), y.max() - y.min(), 1e-12)
    assert tuple(t.root()) == ((x.min() + x.max()) * 0.5, (y.min() + y.max()) * 0.5,
                               span * 0.500000000001 + 1e-12)


# --------------------------------------------------------- f32 traversal
def f32_direct_sorted(state, idx, eps, g):
    """Independent numpy f32 restatement: theta=0 visits every leaf in sorted order."""
    n = len(state)
    xs = state[idx, 0].astype(np.float32)
    ys = state[idx, 1].astype(np.float32)
    ms = state[idx, 2].astype(np.float32)
    e2 = np.float32(eps) * np.float32(eps)
    g = np.float32(g)
    out = np.zeros((n, 2), np.float32)
    for t in range(n):
        tx, ty = np.float32(state[t, 0]), np.float32(state[t, 1])
        mask = idx != t
        dx = xs[mask] - tx
        dy = ys[mask] - ty
        r2 = (dx * dx + dy * dy) + e2
        inv = np.float32(1.0) / np.sqrt(r2)
        sc = ((g * ms[mask]) * inv) * inv * inv
        px, py = dx * sc, dy * sc
        ax = np.float32(0.0)
        ay = np.float32(0.0)
        for a, b in zip(px, py):
            ax = np.float32(ax + a)
            ay = np.float32(ay + b)
        out[t] = (ax, ay)
    return out


def test_f32_traversal_theta0_matches_numpy_restatement(oracle):
    s = ics.rotating_disc(96, 11, softening=0.01)
    cfg = oracle.make_config(theta=0.0, softening=0.01, g=1.0, bucket=2, max_depth=16)
    t = oracle.tree_from_state(s, cfg)
    st, bodies, entries, cells = t.pack()
    acc, stats, hsh = oracle.traverse_f32(st, bodies, entries, cells, evidence=True)
    _, idx, _ = t.entries()
    ref = f32_direct_sorted(s, idx, 0.01, 1.0)
    assert np.array_equal(acc[:, :2], ref)
    assert (stats[:, 2] == 95).all() and (stats[:, 1] == 0).all()
    assert (stats[:, 0] == len(cells)).all()


def test_f32_traversal_close_to_f64_flat(oracle):
    s = ics.rotating_disc(768, 303, softening=0.01)
    cfg = oracle.make_config(theta=0.5, softening=0.01, g=1.0, bucket=1, max_depth=16)
    t = oracle.tree_from_state(s, cfg)
    acc = oracle.traverse_f32(*t.pack())
    ax, ay, _ = t.accelerations()
    assert rel_err(acc[:, 0].astype(np.float64), acc[:, 1].astype(np.float64), ax, ay).max() < 1e-4


def test_interaction_hash_definition(oracle):
    """Per-target hash equals FNV-mix64 over the list rebuilt in Python from a stack DFS."""
    s = ics.uniform_collapse(200, 9)
    cfg = oracle.make_config(theta=0.7, softening=0.005, g=1.0, bucket=4, max_depth=16)
    t = oracle.tree_from_state(s, cfg)
    st, bodies, entries, cells = t.pack()
    _, stats, hsh = oracle.traverse_f32(st, bodies, entries, cells, evidence=True)
    f = cells.view(np.float32)
    M = (1 << 64) - 1

    def mix(h, w):
        h ^= w
        h = (h * 0x100000001b3) & M
        return h ^ (((w << 23) | (w >> 41)) & M)

    for tgt in (0, 57, 199):
        tx, ty = np.float32(s[tgt, 0]), np.float32(s[tgt, 1])
        tpos = bodies[tgt, 4]
        h, stack = 0xcbf29ce484222325, [0]
        while stack:
            c = stack.pop()
            ch = cells[c, 12:16]
            if (ch == NO_CHILD).all():
                for p in range(cells[c, 8], cells[c, 9]):
                    if entries[p, 1] != tgt:
                        h = mix(h, (1 << 60) | int(entries[p, 1]))
                continue
            dx, dy = f[c, 4] - tx, f[c, 5] - ty
            d2 = np.float32(dx * dx) + np.float32(dy * dy)
            inside = cells[c, 8] <= tpos < cells[c, 9]
            if not inside and d2 > 0 and np.float32(2.0) * f[c, 2] / np.sqrt(np.float32(d2)) < np.float32(0.7):
                h = mix(h, (2 << 60) | (int(cells[c, 10]) << 32) | int(cells[c, 8]))
                continue
            stack.extend(int(x) for x in ch[::-1] if x != NO_CHILD)
        assert h == int(hsh[tgt])


# --------------------------------------------------------------- KDK, evidence
def test_kdk_closed_form(oracle):
    rng = np.random.default_rng(4)
    s = rng.normal(0, 1, (50, 8)).astype(np.float32)
    a = rng.normal(0, 1, (50, 4)).astype(np.float32)
    dt = np.float32(1e-3)
    h = np.float32(0.5) * dt
    e = s.copy()
    hv = e[:, 4:6] + h * a[:, 0:2]
    e[:, 0:2] = e[:, 0:2] + dt * hv
    e[:, 4:6] = hv
    got = s.copy()
    oracle.kick_drift_f32(got, a, dt)
    assert np.array_equal(got, e)
    e[:, 4:6] = e[:, 4:6] + h * a[:, 0:2]
    oracle.final_kick_f32(got, a, dt)
    assert np.array_equal(got, e)


def test_sparse_evidence_and_checksum(oracle):
    s = ics.rotating_disc(300, 5, softening=0.01)
    cfg = oracle.make_config(theta=0.5, softening=0.01, g=1.0, bucket=4, max_depth=16)
    t = oracle.tree_from_state(s, cfg)
    meta, entries, cells = t.sparse_evidence()
    n = len(s)
    slots = cells[:, 10].astype(np.int64) * n + cells[:, 8]
    assert (np.diff(slots) > 0).all() and (cells[:, 11] == 1).all()
    for c in cells:
        for q in c[12:16]:
            if q != NO_CHILD:
                assert q % n >= c[8] and q // n == c[10] + 1
    assert oracle.validate_parallel_order(entries)
    # checksum restated in Python
    M = (1 << 64) - 1

    def mix(h, w):
        h ^= w
        h = (h * 0x100000001b3) & M
        return h ^ (((w << 23) | (w >> 41)) & M)

    h = 0xcbf29ce484222325
    for w in list(meta) + [int(v) for v in entries.reshape(-1)] + [int(v) for v in cells.reshape(-1)]:
        h = mix(h, int(w))
    assert h == oracle.parallel_tree_checksum(meta, entries, cells)
    bad = entries.copy()
    bad[[3, 4]] = bad[[4, 3]]
    assert not oracle.validate_parallel_order(bad)


def test_ics_repeatable_and_bounded():
    a = ics.rotating_disc(1000, 303)
    assert np.array_equal(a, ics.rotating_disc(1000, 303))
    assert not np.array_equal(a, ics.rotating_disc(1000, 304))
    r = np.hypot(a[:, 0], a[:, 1])
    assert r.max() <= 5.0 + 1e-6 and abs(a[:, 2].astype(np.float64).sum() - 1.0) < 1e-5
    c = ics.uniform_collapse(1000, 303)
    assert np.hypot(c[:, 0], c[:, 1]).max() <= 1.0 + 1e-6 and (c[:, 4:6] == 0).all()
    b = ics.binary_disc(500, 500)
    m = b[:, 2].astype(np.float64)
    com = (m[:, None] * b[:, [0, 1, 4, 5]].astype(np.float64)).sum(0) / m.sum()
    assert np.abs(com).max() < 0.25  # COM frame, up to the discs' sampling noise (~2/sqrt(500))
