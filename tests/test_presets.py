"""The reference presets make_disc / make_collision (nbody/src/lib.rs:432-528),
restated in gbh_ic_reference_disc / gbh_ic_reference_collision, pinned without
retyping their formulas.

- The generator (lib.rs:432-452, Marsaglia's 64-bit xorshift with shifts 13, 7,
  17, seed max(seed, 1), draws (x >> 11) / 2^53) is rebuilt here as what it is
  algebraically: one 64 x 64 bit matrix over GF(2), the product of the three
  shear maps, applied to the state bit vector.
- Each body's two draws are recovered from its position by inverting the
  geometry (the radius gives the first draw through the Rayleigh quantile, the
  polar angle the second) and compared with the matrix generator's draws, in
  order, for every body.
- The velocity is checked as what it must be physically: perpendicular to the
  radius vector, turning in the disc's direction, and of magnitude spin times the
  circular speed sqrt(G M(<r) / sqrt(r^2 + eps^2)) of an exponential disc, whose
  enclosed mass fraction is integrated numerically here (not taken from its
  closed form); G is checked against its physical value in kpc^3 / (Msun Myr^2).

The transcendental functions (ln, sqrt, sin, cos, exp) are the platform libm's on
both sides, so agreement is to rounding, not bit for bit (which the reference,
holding no preset vector, does not fix).
"""
import math

import numpy as np
from scipy.integrate import quad

from galaxy_amd import ics

G_REF = 4.498502151575286e-12  # nbody/src/lib.rs:6
SOFTENING = 0.05               # nbody/src/lib.rs:8 (DEFAULT_SOFTENING_KPC)


def _shift_matrix(s):
    """x << s (s > 0) or x >> -s as a 64 x 64 GF(2) matrix acting on bit vectors (bit i = row i)."""
    m = np.zeros((64, 64), np.uint8)
    for i in range(64):
        j = i + s
        if 0 <= j < 64:
            m[j, i] = 1
    return m


def _xorshift_matrix():
    eye = np.eye(64, dtype=np.uint8)
    a = (eye + _shift_matrix(13)) % 2   # x ^= x << 13
    b = (eye + _shift_matrix(-7)) % 2   # x ^= x >> 7
    c = (eye + _shift_matrix(17)) % 2   # x ^= x << 17
    return (c.astype(np.int64) @ ((b.astype(np.int64) @ a.astype(np.int64)) % 2)) % 2


def _bits(x):
    return np.array([(x >> i) & 1 for i in range(64)], np.int64)


def _word(v):
    return sum(int(b) << i for i, b in enumerate(v))


def matrix_draws(seed, count):
    """`count` unit draws of the generator seeded with `seed`, by the GF(2) matrix."""
    m = _xorshift_matrix()
    s = _bits(max(seed, 1))
    out = []
    for _ in range(count):
        s = (m @ s) % 2
        out.append((_word(s) >> 11) / float(1 << 53))
    return np.array(out)


def test_matrix_generator_is_marsaglia_xorshift64():
    # the matrix is invertible (a full-period xorshift is a bijection on nonzero states)
    # and its first outputs from seed 1 are nonzero and distinct
    m = _xorshift_matrix()
    s = _bits(1)
    seen = set()
    for _ in range(64):
        s = (m @ s) % 2
        w = _word(s)
        assert w != 0 and w not in seen
        seen.add(w)
    # seed 0 is replaced by 1 (Self { state: seed.max(1) })
    assert np.array_equal(matrix_draws(0, 4), matrix_draws(1, 4))


def enclosed_fraction(q):
    """M(<r)/M of an exponential surface density exp(-R/h), q = r / h, by quadrature."""
    if q <= 0.0:
        return 0.0
    v, _ = quad(lambda t: t * math.exp(-t), 0.0, q, epsabs=1e-15, epsrel=1e-13, limit=200)
    return min(max(v, 0.0), 1.0)


def check_disc(bodies, draws, scale, total_mass, per_body_mass, center, cvel, direction, spin):
    cx, cy = center
    n = len(bodies)
    assert np.all(bodies[:, 4] == per_body_mass)
    dx, dy = bodies[:, 0] - cx, bodies[:, 1] - cy
    r = np.hypot(dx, dy)
    # first draw: r = scale sqrt(-ln(1 - u1))  =>  u1 = 1 - exp(-(r / scale)^2)
    u1 = -np.expm1(-(r / scale) ** 2)
    assert np.max(np.abs(u1 - draws[0::2][:n])) < 1e-9
    # second draw: polar angle = 2 pi u2
    u2 = np.mod(np.arctan2(dy, dx), 2.0 * math.pi) / (2.0 * math.pi)
    d = np.abs(u2 - draws[1::2][:n])
    assert np.max(np.minimum(d, 1.0 - d)) < 1e-9
    vx, vy = bodies[:, 2] - cvel[0], bodies[:, 3] - cvel[1]
    speed = np.hypot(vx, vy)
    # perpendicular to the radius, turning in the disc's direction
    assert np.max(np.abs(vx * dx + vy * dy) / np.maximum(speed * r, 1e-300)) < 1e-9
    assert np.all(np.sign(dx * vy - dy * vx) == direction)
    # magnitude: spin times the circular speed of an exponential disc.  The enclosed
    # fraction the speed implies, F = (speed / spin)^2 sqrt(r^2 + eps^2) / (G M), equals
    # the quadrature's to 1e-9 relative, plus a few units of 2^-53 absolute: the
    # reference's closed form 1 - e^-q (1 + q) cancels for small q.
    soft = np.maximum(np.sqrt(r * r + SOFTENING * SOFTENING), 1e-9)
    implied = (speed / spin) ** 2 * soft / (G_REF * total_mass)
    exact = np.array([enclosed_fraction(ri / scale) for ri in r])
    assert np.all(np.abs(implied - exact) <= 1e-9 * exact + 1e-15)


def test_gravitational_constant_units():
    # G = 4.30091727e-6 kpc (km/s)^2 / Msun; 1 km/s = 1.0227121650537077e-3 kpc/Myr
    g = 4.30091727e-6 * 1.0227121650537077e-3 ** 2
    assert abs(g - G_REF) / G_REF < 1e-6


def test_make_disc_draws_and_velocities():
    for n, seed, mass, scale in [(512, 303, 5.0e10, 3.0), (777, 0, 1.0e9, 1.5), (64, 2**63 + 5, 3.0e10, 4.0)]:
        b = ics.reference_disc(n, seed, mass, scale)
        draws = matrix_draws(seed, 2 * n)
        check_disc(b, draws, scale, mass, mass / n, (0.0, 0.0), (0.0, 0.0), 1.0, 0.88)


def test_make_collision_two_discs():
    for n, seed in [(1500, 303), (257, 2**64 - 1)]:
        b = ics.reference_collision(n, seed)
        left, right = n // 2, n - n // 2
        half = 5.0e10 * 0.5
        check_disc(b[:left], matrix_draws(seed, 2 * left), 3.0, half, half / left, (-7.5, -1.2),
                   (0.030, 0.010), 1.0, 0.86)
        # the second disc's generator is seeded with seed + 1, wrapping at 2^64
        check_disc(b[left:], matrix_draws((seed + 1) % 2**64, 2 * right), 2.7, half, half / right,
                   (7.5, 1.2), (-0.030, -0.010), -1.0, 0.86)
