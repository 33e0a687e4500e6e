"""Pins of the oracle's error metrics and FP64 diagnostics.

* direct_for: direct_acceleration_for (nbody/src/lib.rs:315-329) written out in
  Python in the reference's operation order (add_point_mass, lib.rs:247-255), equal
  bit for bit, and the two-body closed form.
* direct_accelerations (lib.rs:331-354): the pairwise form equals the per-target sums
  to rounding, and Newton's third law makes the mass-weighted sum vanish.
* probe_error (lib.rs:370-395): the probe index rule k*(N-1)/(P-1), and the relative
  error of a known perturbation.
* force_error / state_error (galaxy-bh-gpu-tree-parallel.rs:62-114): closed forms of
  uniform perturbations and the floors of the denominators.
* energy (kinetic + Plummer-softened pair potential, Lz): a two-body closed form and a
  numpy brute force; this diagnostic has no counterpart in the reference.
"""
import math

import numpy as np
import pytest


def cfg_of(oracle, eps=0.05, g=1.0):
    return oracle.make_config(theta=0.5, softening=eps, g=g, bucket=4, max_depth=16)


def bodies(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    y = rng.normal(size=n)
    m = rng.uniform(0.5, 2.0, size=n)
    vx = rng.normal(size=n)
    vy = rng.normal(size=n)
    return x, y, m, vx, vy


def direct_python(x, y, m, i, eps, g):
    ax = ay = 0.0
    for j in range(len(x)):
        if j == i:
            continue
        dx = float(x[j]) - float(x[i])
        dy = float(y[j]) - float(y[i])
        r2 = dx * dx + dy * dy + eps * eps
        inv_r = 1.0 / math.sqrt(r2)
        scale = g * float(m[j]) * inv_r * inv_r * inv_r
        ax += dx * scale
        ay += dy * scale
    return ax, ay


def test_direct_for_equals_reference_order(oracle):
    x, y, m, _, _ = bodies(37, 1)
    eps, g = 0.05, 2.5
    idx = np.array([0, 5, 36], np.uint32)
    ex, ey = oracle.direct_for(x, y, m, cfg_of(oracle, eps, g), idx)
    for k, i in enumerate(idx):
        px, py = direct_python(x, y, m, int(i), eps, g)
        assert ex[k] == px and ey[k] == py


def test_direct_two_body_closed_form(oracle):
    d, eps, g, m1 = 3.0, 0.4, 1.7, 2.0
    x = np.array([0.0, d])
    y = np.array([0.0, 0.0])
    m = np.array([1.0, m1])
    ex, ey = oracle.direct_for(x, y, m, cfg_of(oracle, eps, g), np.array([0, 1], np.uint32))
    want = g * m1 * d / (d * d + eps * eps) ** 1.5
    assert ex[0] == pytest.approx(want, rel=1e-14) and ey[0] == 0.0
    assert ex[1] == pytest.approx(-g * 1.0 * d / (d * d + eps * eps) ** 1.5, rel=1e-14)


def test_pairwise_direct_matches_per_target_and_newton(oracle):
    x, y, m, _, _ = bodies(60, 2)
    cfg = cfg_of(oracle, 0.1, 1.0)
    px, py = oracle.direct_accelerations(x, y, m, cfg)
    ex, ey = oracle.direct_for(x, y, m, cfg, np.arange(60, dtype=np.uint32))
    scale = np.max(np.hypot(ex, ey))
    assert np.max(np.abs(px - ex)) < 1e-13 * scale and np.max(np.abs(py - ey)) < 1e-13 * scale
    assert abs(np.sum(m * px)) < 1e-12 * scale * m.sum() and abs(np.sum(m * py)) < 1e-12 * scale * m.sum()


def test_probe_error_index_rule_and_known_perturbation(oracle):
    n, probes = 50, 8
    x, y, m, _, _ = bodies(n, 3)
    cfg = cfg_of(oracle, 0.05, 1.0)
    exact = oracle.direct_for(x, y, m, cfg, np.arange(n, dtype=np.uint32))
    ax, ay = np.array(exact[0]), np.array(exact[1])
    # perturb one probe body (k = 3: index 3 * 49 // 7 = 21) by a known relative amount
    probe_idx = [k * (n - 1) // (probes - 1) for k in range(probes)]
    assert probe_idx == [0, 7, 14, 21, 28, 35, 42, 49]
    delta = 1e-3
    bx, by = ax.copy(), ay.copy()
    bx[21] = ax[21] * (1 + delta)
    by[21] = ay[21] * (1 + delta)
    # perturbing a non-probe body changes nothing
    bx[22] = ax[22] * 5.0
    by[22] = ay[22] * 5.0
    e = oracle.probe_error(x, y, m, bx, by, cfg, probes)
    assert e["max_relative"] == pytest.approx(delta, rel=1e-9)
    assert e["rms_relative"] == pytest.approx(delta / math.sqrt(probes), rel=1e-9)
    # probes clamp to [1, N]; one probe is index 0
    e1 = oracle.probe_error(x, y, m, bx, by, cfg, 1)
    assert e1["max_relative"] == 0.0
    ea = oracle.probe_error(x, y, m, bx, by, cfg, 10 * n)
    assert ea["max_relative"] == pytest.approx(4.0, rel=1e-12)


def test_force_error_closed_forms(oracle):
    rng = np.random.default_rng(4)
    bx, by = rng.normal(size=100), rng.normal(size=100)
    for d in (0.0, 1e-6, 0.25):
        e = oracle.force_error(bx * (1 + d), by * (1 + d), bx, by)
        # b * (1 + d) - b carries a rounding error of order 1e-16 / d relative to d
        assert e["rms_relative"] == pytest.approx(d, rel=1e-8, abs=1e-300)
        assert e["max_relative"] == pytest.approx(d, rel=1e-8, abs=1e-300)
    # one body off by a rotation of 90 degrees: relative error sqrt(2) there
    ax, ay = bx.copy(), by.copy()
    ax[7], ay[7] = -by[7], bx[7]
    e = oracle.force_error(ax, ay, bx, by)
    assert e["max_relative"] == pytest.approx(math.sqrt(2), rel=1e-12)
    assert e["rms_relative"] == pytest.approx(math.sqrt(2 / 100), rel=1e-12)
    # zero reference: the denominator floors at 1e-30
    e = oracle.force_error(np.array([1e-31]), np.array([0.0]), np.array([0.0]), np.array([0.0]))
    assert e["max_relative"] == pytest.approx(0.1, rel=1e-12)


def test_state_error_closed_forms(oracle):
    rng = np.random.default_rng(5)
    b = [rng.normal(size=64) + 3.0 for _ in range(4)]
    d = 1e-4
    a = [b[0] * (1 + d), b[1] * (1 + d), b[2] * (1 - 2 * d), b[3] * (1 - 2 * d)]
    e = oracle.state_error(a, b)
    assert e["position_rms"] == pytest.approx(d, rel=1e-9)
    assert e["position_max"] == pytest.approx(d, rel=1e-9)
    assert e["velocity_rms"] == pytest.approx(2 * d, rel=1e-9)
    assert e["velocity_max"] == pytest.approx(2 * d, rel=1e-9)
    # per-body floors: positions at the origin divide by 1e-6, velocities at rest by 1e-9
    z = np.zeros(1)
    e = oracle.state_error((np.array([1e-8]), z, np.array([1e-12]), z), (z, z, z, z))
    assert e["position_max"] == pytest.approx(1e-2, rel=1e-12)
    assert e["velocity_max"] == pytest.approx(1e-3, rel=1e-12)


def test_energy_two_body_closed_form(oracle):
    d, eps, g = 2.0, 0.3, 1.5
    x = np.array([0.0, d])
    y = np.array([0.0, 0.0])
    m = np.array([1.0, 3.0])
    vx = np.array([0.0, 0.5])
    vy = np.array([1.0, -0.25])
    e = oracle.energy(x, y, vx, vy, m, eps, g)
    kin = 0.5 * 1.0 * 1.0 + 0.5 * 3.0 * (0.25 + 0.0625)
    pot = -g * 1.0 * 3.0 / math.sqrt(d * d + eps * eps)
    assert e["kinetic"] == pytest.approx(kin, rel=1e-15)
    assert e["energy"] == pytest.approx(kin + pot, rel=1e-14)
    assert e["lz"] == pytest.approx(1.0 * (0.0 * 1.0) + 3.0 * (d * -0.25 - 0.0 * 0.5), rel=1e-15)


def test_energy_numpy_brute_force(oracle):
    x, y, m, vx, vy = bodies(300, 6)
    eps, g = 0.02, 0.7
    e = oracle.energy(x, y, vx, vy, m, eps, g)
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    inv = 1.0 / np.sqrt(dx * dx + dy * dy + eps * eps)
    iu = np.triu_indices(len(x), 1)
    pot = -g * np.sum((m[:, None] * m[None, :] * inv)[iu])
    kin = 0.5 * np.sum(m * (vx * vx + vy * vy))
    lz = np.sum(m * (x * vy - y * vx))
    assert e["kinetic"] == pytest.approx(kin, rel=1e-13)
    assert e["energy"] == pytest.approx(kin + pot, rel=1e-12)
    assert e["lz"] == pytest.approx(lz, rel=1e-12, abs=1e-12)
