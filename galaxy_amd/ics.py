"""Seeded synthetic workloads (host side, through the library's IC generators).

The five measured configurations:
  C0  rotating exponential disc, N=768, seed 303, G=M=1, eps=0.01, dt=1e-3, 200 steps, bucket 1
  C1  cold collapse, uniform unit disc, N=65,536, seed 303, eps=0.005, dt=5e-4, 1000 steps, bucket 4
  C2  binary disc encounter, 2 x 524,288 bodies, masses 1 : 0.5, eps=0.002, theta 0.6, dt=2.5e-4
  C3  256 rotating discs, N log-uniform in [512, 65536], scale length U[0.5, 2], seeds 1000+i
  C4  serving stream of 20,000 'advance k steps' requests (config4_requests)
Every generator returns frozen GpuEvolveState records, shape (n, 8) float32.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .bh import BhError, Config


def _check(rc):
    if rc != 0:
        raise BhError(L.lib().gbh_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def rotating_disc(n, seed, total_mass=1.0, scale=1.0, truncation=5.0, softening=0.01, g=1.0,
                  dispersion=0.05) -> np.ndarray:
    out = np.zeros((n, 8), np.float32)
    _check(L.lib().gbh_ic_rotating_disc(n, seed, total_mass, scale, truncation, softening, g,
                                        dispersion, _p(out)))
    return out


def uniform_collapse(n, seed, total_mass=1.0, radius=1.0) -> np.ndarray:
    out = np.zeros((n, 8), np.float32)
    _check(L.lib().gbh_ic_uniform_collapse(n, seed, total_mass, radius, _p(out)))
    return out


def binary_disc(n1, n2, seed1=303, seed2=304, m1=1.0, m2=0.5, separation=8.0, pericentre=1.0,
                softening=0.002, g=1.0, dispersion=0.05) -> np.ndarray:
    out = np.zeros((n1 + n2, 8), np.float32)
    _check(L.lib().gbh_ic_binary_disc(n1, n2, seed1, seed2, m1, m2, separation, pericentre,
                                      softening, g, dispersion, _p(out)))
    return out


def reference_disc(n, seed=303, total_mass=5.0e10, scale_kpc=3.0) -> np.ndarray:
    """make_disc (nbody/src/lib.rs:485-502) -> f64 (n, 5) {x, y, vx, vy, mass}."""
    out = np.zeros((n, 5), np.float64)
    _check(L.lib().gbh_ic_reference_disc(n, seed, total_mass, scale_kpc, _p(out)))
    return out


def reference_collision(n, seed=303) -> np.ndarray:
    """make_collision (nbody/src/lib.rs:504-528) -> f64 (n, 5)."""
    out = np.zeros((n, 5), np.float64)
    _check(L.lib().gbh_ic_reference_collision(n, seed, _p(out)))
    return out


def bodies_to_state(b: np.ndarray) -> np.ndarray:
    """f64 bodies {x, y, vx, vy, mass} -> GpuEvolveState f32 (create_evolving_state packing)."""
    s = np.zeros((len(b), 8), np.float32)
    s[:, 0] = b[:, 0]
    s[:, 1] = b[:, 1]
    s[:, 2] = b[:, 4]
    s[:, 4] = b[:, 2]
    s[:, 5] = b[:, 3]
    return s


@dataclass
class Workload:
    name: str
    states: list            # list of (n, 8) f32 arrays, one per simulation
    config: Config
    dt: float
    steps: int
    extra: dict = field(default_factory=dict)

    @property
    def bodies(self) -> int:
        return int(sum(len(s) for s in self.states))


class _XorShift64:
    """XorShift64 (nbody/src/lib.rs:430-452) for ensemble-level draws."""

    def __init__(self, seed):
        self.s = seed if seed else 1

    def unit(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = x
        return (x >> 11) / float(1 << 53)


def config0(n=768, steps=200) -> Workload:
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=1, max_depth=16)
    return Workload("C0 rotating disc N=768", [rotating_disc(n, 303, softening=0.01)], cfg, 1e-3, steps)


def config1(n=65536, steps=1000) -> Workload:
    cfg = Config(theta=0.5, softening_kpc=0.005, g=1.0, bucket=4, max_depth=16)
    return Workload("C1 cold collapse N=65536", [uniform_collapse(n, 303)], cfg, 5e-4, steps)


def config2(n_each=524288, steps=500) -> Workload:
    cfg = Config(theta=0.6, softening_kpc=0.002, g=1.0, bucket=4, max_depth=16)
    return Workload("C2 binary disc N=1048576", [binary_disc(n_each, n_each)], cfg, 2.5e-4, steps)


def ensemble_sizes(count=256, lo=512, hi=65536, seed=20260):
    rng = _XorShift64(seed)
    sizes, scales = [], []
    for _ in range(count):
        u = rng.unit()
        sizes.append(int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))))))
        scales.append(0.5 + 1.5 * rng.unit())
    return sizes, scales


def config3(count=256, steps=500, bucket=4) -> Workload:
    sizes, scales = ensemble_sizes(count)
    states = [rotating_disc(n, 1000 + i, scale=sc, truncation=5.0, softening=0.01)
              for i, (n, sc) in enumerate(zip(sizes, scales))]
    cfg = Config(theta=0.5, softening_kpc=0.01, g=1.0, bucket=bucket, max_depth=16)
    return Workload(f"C3 ensemble {count} discs", states, cfg, 1e-3, steps,
                    extra={"sizes": sizes, "scales": scales})


def config4_requests(count=20000, rate_lo=50.0, rate_hi=2000.0, seed=777):
    """The C4 request stream: Poisson arrivals whose rate ramps linearly from
    rate_lo to rate_hi over count / mean-rate seconds (19.5 s for 20,000), N
    log-uniform in [1,024, 262,144], k uniform in {10, 100, 1,000}, initial
    conditions alternating between a rotating disc (C0 physics: eps 0.01, dt 1e-3,
    leaf size 1) and a cold collapse (C1 physics: eps 0.005, dt 5e-4, leaf size 4),
    seed 10^6 + request index.  Returns [(arrival s, kind, n, seed, steps)]; the
    states come from request_state."""
    rng = np.random.default_rng(seed)
    span = count / ((rate_lo + rate_hi) / 2.0)
    t, out = 0.0, []
    for k in range(count):
        rate = rate_lo + (rate_hi - rate_lo) * min(t / span, 1.0)
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(1024), math.log(262144)))))
        steps = int(rng.choice([10, 100, 1000]))
        out.append((t, "disc" if k % 2 == 0 else "collapse", n, 10**6 + k, steps))
    return out


def request_physics(kind):
    """(Config, dt) of a C4 request of the given kind ("disc": C0's, "collapse": C1's)."""
    w = config0(n=2, steps=0) if kind == "disc" else config1(n=2, steps=0)
    return w.config, w.dt


def request_state(kind, n, seed) -> np.ndarray:
    """Initial state of a C4 request."""
    if kind == "disc":
        return rotating_disc(n, seed, softening=request_physics("disc")[0].softening_kpc)
    return uniform_collapse(n, seed)
