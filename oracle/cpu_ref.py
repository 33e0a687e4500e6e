"""ctypes wrapper for the CPU oracle (oracle/libcpu_ref.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py import this
module; the galaxy_amd package never does.

Record arrays use the frozen transfer ABI of runtime/src/nbody_gpu.rs:22-83:
  state  (n, 8) float32  = GpuEvolveState {x, y, m, 0, vx, vy, 0, 0}
  accel  (n, 4) float32  = GpuAccel {ax, ay, 0, 0}
  entry  (n, 4) uint32   = GpuEntry {code, body, 0, 0}
  cell   (k, 16) uint32  = GpuCell (floats viewed as uint32 bit patterns)
  body   (n, 8) uint32   = GpuBody {x, y, m, 0 | position, 0, 0, 0}
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GALAXY_CPU_REF_LIB selects another build of the oracle (the sanitized one of
# `make -C oracle sanitize`).
_LIB_PATH = os.environ.get("GALAXY_CPU_REF_LIB") or os.path.join(_HERE, "libcpu_ref.so")
_lib = None


class Config(C.Structure):
    """galaxy_nbody::Config (nbody/src/lib.rs:27-46)."""

    _fields_ = [
        ("theta", C.c_double),
        ("softening", C.c_double),
        ("g", C.c_double),
        ("bucket", C.c_uint32),
        ("max_depth", C.c_uint32),
    ]


def make_config(theta=0.5, softening=0.05, g=4.498502151575286e-12, bucket=4, max_depth=48):
    return Config(theta, softening, g, bucket, max_depth)


class OracleError(RuntimeError):
    pass


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        vp, u32, u64, dbl, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
        cfgp = C.POINTER(Config)
        sig = {
            "cref_last_error": ([], C.c_char_p),
            "cref_config_validate": ([cfgp], i32),
            "cref_morton2": ([u32, u32], u32),
            "cref_quantize_axis": ([dbl, dbl, dbl], u32),
            "cref_build_flat_tree": ([u32, vp, vp, vp, cfgp], vp),
            "cref_tree_free": ([vp], None),
            "cref_tree_counts": ([vp, vp], None),
            "cref_tree_root": ([vp, vp], None),
            "cref_tree_entries": ([vp, vp, vp, vp], None),
            "cref_tree_cells": ([vp, vp, vp], None),
            "cref_flat_topology_checksum": ([vp], u64),
            "cref_pack_flat_tree": ([vp, u32, vp, vp, vp, cfgp, vp, vp, vp, vp], i32),
            "cref_sparse_evidence": ([vp, cfgp, vp, vp, vp], None),
            "cref_parallel_tree_checksum": ([vp, u32, vp, u32, vp], u64),
            "cref_validate_parallel_order": ([u32, vp], i32),
            "cref_traverse_f32": ([vp, vp, vp, vp, vp, vp, vp], i32),
            "cref_set_threads": ([i32], None),
            "cref_flat_accelerations": ([vp, u32, vp, vp, vp, cfgp, vp, vp, vp], i32),
            "cref_recursive_accelerations": ([u32, vp, vp, vp, cfgp, vp, vp, vp, vp], i32),
            "cref_direct_accelerations": ([u32, vp, vp, vp, cfgp, vp, vp], i32),
            "cref_direct_for": ([u32, vp, vp, vp, cfgp, u32, vp, vp, vp, i32], i32),
            "cref_probe_error": ([u32, vp, vp, vp, vp, vp, cfgp, u32, vp, i32], i32),
            "cref_energy": ([u32, vp, vp, vp, vp, vp, dbl, dbl, vp, i32], None),
            "cref_kick_drift_f32": ([u32, vp, vp, C.c_float], None),
            "cref_final_kick_f32": ([u32, vp, vp, C.c_float], None),
            "cref_flat_evolve": ([u32, vp, vp, vp, vp, vp, u32, dbl, cfgp], i32),
            "cref_f32_evolve": ([u32, vp, vp, u32, dbl, cfgp, vp, vp], i32),
            "cref_state_checksum": ([u32, vp, vp, vp, vp, vp], u64),
            "cref_force_error": ([u32, vp, vp, vp, vp, vp], None),
            "cref_state_error": ([u32, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
            "cref_gpu_tree_build": ([u32, vp, u32, u32, vp, vp, vp, u32, vp, u32], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = res
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc):
    if rc != 0:
        raise OracleError(lib().cref_last_error().decode())


def threads_default():
    return max(1, min(16, os.cpu_count() or 1))


def set_threads(threads):
    """Threads of the f32 traversal (every count gives identical results)."""
    lib().cref_set_threads(int(threads))


class FlatTree:
    """BH #2A flat tree (nbody/src/flat.rs:217-268), built from f64 bodies."""

    def __init__(self, x, y, m, cfg: Config):
        self.x, self.y, self.m = _f64(x), _f64(y), _f64(m)
        self.n = len(self.x)
        self.cfg = cfg
        h = lib().cref_build_flat_tree(self.n, _p(self.x), _p(self.y), _p(self.m), C.byref(cfg))
        if not h:
            raise OracleError(lib().cref_last_error().decode())
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.cref_tree_free(h)
            self._h = None

    def counts(self):
        c = np.zeros(4, np.uint32)
        lib().cref_tree_counts(self._h, _p(c))
        return dict(cells=int(c[0]), node_count=int(c[1]), leaf_count=int(c[2]), max_depth=int(c[3]))

    def root(self):
        r = np.zeros(3, np.float64)
        lib().cref_tree_root(self._h, _p(r))
        return r

    def entries(self):
        code = np.zeros(self.n, np.uint32)
        idx = np.zeros(self.n, np.uint32)
        pos = np.zeros(self.n, np.uint32)
        lib().cref_tree_entries(self._h, _p(code), _p(idx), _p(pos))
        return code, idx, pos

    def cells(self):
        k = self.counts()["cells"]
        geom = np.zeros((k, 6), np.float64)
        topo = np.zeros((k, 7), np.uint32)
        lib().cref_tree_cells(self._h, _p(geom), _p(topo))
        return geom, topo

    def topology_checksum(self):
        return int(lib().cref_flat_topology_checksum(self._h))

    def pack(self):
        """pack_flat_tree -> (settings u32[8], bodies u32[n,8], entries u32[n,4], cells u32[k,16])."""
        k = self.counts()["cells"]
        st = np.zeros(8, np.uint32)
        bodies = np.zeros((self.n, 8), np.uint32)
        entries = np.zeros((self.n, 4), np.uint32)
        cells = np.zeros((k, 16), np.uint32)
        _check(lib().cref_pack_flat_tree(self._h, self.n, _p(self.x), _p(self.y), _p(self.m),
                                         C.byref(self.cfg), _p(st), _p(bodies), _p(entries), _p(cells)))
        return st, bodies, entries, cells

    def sparse_evidence(self):
        """BH #2D-layout evidence -> (meta u32[8], entries u32[n,4], cells u32[k,16])."""
        k = self.counts()["cells"]
        meta = np.zeros(8, np.uint32)
        entries = np.zeros((self.n, 4), np.uint32)
        cells = np.zeros((k, 16), np.uint32)
        lib().cref_sparse_evidence(self._h, C.byref(self.cfg), _p(meta), _p(entries), _p(cells))
        return meta, entries, cells

    def accelerations(self):
        ax = np.zeros(self.n)
        ay = np.zeros(self.n)
        st = np.zeros(3, np.uint64)
        _check(lib().cref_flat_accelerations(self._h, self.n, _p(self.x), _p(self.y), _p(self.m),
                                             C.byref(self.cfg), _p(ax), _p(ay), _p(st)))
        return ax, ay, dict(visited=int(st[0]), approximated=int(st[1]), direct=int(st[2]))


def parallel_tree_checksum(meta, entries, cells):
    meta = np.ascontiguousarray(meta, np.uint32)
    entries = np.ascontiguousarray(entries, np.uint32)
    cells = np.ascontiguousarray(cells, np.uint32)
    return int(lib().cref_parallel_tree_checksum(_p(meta), len(entries), _p(entries), len(cells), _p(cells)))


def validate_parallel_order(entries):
    entries = np.ascontiguousarray(entries, np.uint32)
    return lib().cref_validate_parallel_order(len(entries), _p(entries)) == 0


def traverse_f32(settings, bodies, entries, cells, evidence=False):
    """nbody.wgsl bh_traverse on packed records -> accel (n,4) f32 [, stats (n,3), hash (n,)]."""
    settings = np.ascontiguousarray(settings, np.uint32)
    bodies = np.ascontiguousarray(bodies, np.uint32)
    entries = np.ascontiguousarray(entries, np.uint32)
    cells = np.ascontiguousarray(cells, np.uint32)
    n = int(settings[0])
    out = np.zeros((n, 4), np.float32)
    stats = np.zeros((n, 3), np.uint32) if evidence else None
    hsh = np.zeros(n, np.uint64) if evidence else None
    _check(lib().cref_traverse_f32(_p(settings), _p(bodies), _p(entries), _p(cells), _p(out),
                                   _p(stats), _p(hsh)))
    return (out, stats, hsh) if evidence else out


def recursive_accelerations(x, y, m, cfg):
    x, y, m = _f64(x), _f64(y), _f64(m)
    n = len(x)
    ax, ay = np.zeros(n), np.zeros(n)
    st = np.zeros(3, np.uint64)
    ts = np.zeros(3, np.uint32)
    _check(lib().cref_recursive_accelerations(n, _p(x), _p(y), _p(m), C.byref(cfg), _p(ax), _p(ay),
                                              _p(st), _p(ts)))
    return ax, ay, dict(node_count=int(ts[0]), leaf_count=int(ts[1]), max_depth=int(ts[2]),
                        visited=int(st[0]), approximated=int(st[1]), direct=int(st[2]))


def direct_accelerations(x, y, m, cfg):
    x, y, m = _f64(x), _f64(y), _f64(m)
    n = len(x)
    ax, ay = np.zeros(n), np.zeros(n)
    _check(lib().cref_direct_accelerations(n, _p(x), _p(y), _p(m), C.byref(cfg), _p(ax), _p(ay)))
    return ax, ay


def direct_for(x, y, m, cfg, index, threads=None):
    x, y, m = _f64(x), _f64(y), _f64(m)
    index = np.ascontiguousarray(index, np.uint32)
    ax, ay = np.zeros(len(index)), np.zeros(len(index))
    _check(lib().cref_direct_for(len(x), _p(x), _p(y), _p(m), C.byref(cfg), len(index), _p(index),
                                 _p(ax), _p(ay), threads or threads_default()))
    return ax, ay


def probe_error(x, y, m, ax, ay, cfg, probes, threads=None):
    x, y, m, ax, ay = map(_f64, (x, y, m, ax, ay))
    out = np.zeros(2)
    _check(lib().cref_probe_error(len(x), _p(x), _p(y), _p(m), _p(ax), _p(ay), C.byref(cfg), probes,
                                  _p(out), threads or threads_default()))
    return dict(rms_relative=float(out[0]), max_relative=float(out[1]))


def energy(x, y, vx, vy, m, softening, g, threads=None):
    x, y, vx, vy, m = map(_f64, (x, y, vx, vy, m))
    out = np.zeros(3)
    lib().cref_energy(len(x), _p(x), _p(y), _p(vx), _p(vy), _p(m), softening, g, _p(out),
                      threads or threads_default())
    return dict(energy=float(out[0]), lz=float(out[1]), kinetic=float(out[2]))


def kick_drift_f32(state, accel, dt):
    lib().cref_kick_drift_f32(len(state), _p(state), _p(np.ascontiguousarray(accel, np.float32)),
                              C.c_float(np.float32(dt)))


def final_kick_f32(state, accel, dt):
    lib().cref_final_kick_f32(len(state), _p(state), _p(np.ascontiguousarray(accel, np.float32)),
                              C.c_float(np.float32(dt)))


def flat_evolve(x, y, vx, vy, m, steps, dt, cfg):
    """cpu_flat_evolve (f64 BH #2A trajectory); returns new (x, y, vx, vy)."""
    x, y, vx, vy, m = [np.array(a, dtype=np.float64) for a in (x, y, vx, vy, m)]
    _check(lib().cref_flat_evolve(len(x), _p(x), _p(y), _p(vx), _p(vy), _p(m), steps, dt, C.byref(cfg)))
    return x, y, vx, vy


def f32_evolve(state, steps, dt, cfg, evidence=False):
    """Drop-in f32 trajectory (2A-exact build + f32 traversal + f32 KDK) on a state copy."""
    s = np.array(state, dtype=np.float32, order="C")
    n = len(s)
    a = np.zeros((n, 4), np.float32)
    stats = np.zeros((n, 3), np.uint32) if evidence else None
    hsh = np.zeros(n, np.uint64) if evidence else None
    _check(lib().cref_f32_evolve(n, _p(s), _p(a), steps, dt, C.byref(cfg), _p(stats), _p(hsh)))
    return (s, a, stats, hsh) if evidence else (s, a)


def state_checksum(x, y, vx, vy, m):
    x, y, vx, vy, m = map(_f64, (x, y, vx, vy, m))
    return int(lib().cref_state_checksum(len(x), _p(x), _p(y), _p(vx), _p(vy), _p(m)))


def force_error(ax, ay, bx, by):
    ax, ay, bx, by = map(_f64, (ax, ay, bx, by))
    out = np.zeros(2)
    lib().cref_force_error(len(ax), _p(ax), _p(ay), _p(bx), _p(by), _p(out))
    return dict(rms_relative=float(out[0]), max_relative=float(out[1]))


def state_error(a, b):
    """a, b: dicts/tuples (x, y, vx, vy) of f64 arrays -> Error4 (galaxy-bh-gpu-tree-parallel.rs:62-90)."""
    a = [_f64(v) for v in a]
    b = [_f64(v) for v in b]
    out = np.zeros(4)
    lib().cref_state_error(len(a[0]), *[_p(v) for v in a], *[_p(v) for v in b], _p(out))
    return dict(position_rms=float(out[0]), position_max=float(out[1]),
                velocity_rms=float(out[2]), velocity_max=float(out[3]))


def tree_from_state(state, cfg):
    """2A tree of the widened f32 state (the build every drop-in force state uses)."""
    s = np.asarray(state, dtype=np.float32)
    return FlatTree(s[:, 0].astype(np.float64), s[:, 1].astype(np.float64),
                    s[:, 2].astype(np.float64), cfg)


class GpuTreeBuild:
    """BH #2C tree (runtime/src/nbody_tree.wgsl:77-370) of an f32 state: meta
    (data u32[4], bounds f32[4]), bodies u32[n, 8], entries u32[n, 4] and the
    level-order cells u32[cells, 16], in the ABI's words."""

    def __init__(self, state, bucket=4, max_depth=48):
        s = np.ascontiguousarray(state, np.float32)
        n = len(s)
        padded = 1 << max(0, (n - 1).bit_length())
        cap = 16 * n + 1
        self.meta = np.zeros(8, np.uint32)
        self.bodies = np.zeros((n, 8), np.uint32)
        entries = np.zeros((padded, 4), np.uint32)
        cells = np.zeros((cap, 16), np.uint32)
        _check(lib().cref_gpu_tree_build(n, _p(s), bucket, max_depth, _p(self.meta), _p(self.bodies),
                                         _p(entries), padded, _p(cells), cap))
        self.n = n
        self.entries = entries[:n]
        self.cells = cells[:int(self.meta[0])] if self.meta[3] == 0 else cells[:0]

    def bounds(self):
        return self.meta[4:8].view(np.float32)

    def settings(self, theta, softening, g):
        """BhSettings of force_from_gpu_tree (nbody_gpu.rs:1303-1325): root cell 0."""
        st = np.zeros(8, np.uint32)
        st[0], st[1], st[2] = self.n, len(self.cells), 0
        st[4:7] = np.array([theta, softening, g], np.float32).view(np.uint32)
        return st

    def accelerations(self, theta, softening, g):
        return traverse_f32(self.settings(theta, softening, g), self.bodies, self.entries, self.cells)


def gpu_tree_evolve(state, steps, dt, theta, softening, g, bucket=4, max_depth=48):
    """run_serial_oracle (galaxy-bh-gpu-tree-parallel.rs:225-246): BH #2C build and
    force before the loop, then kick/drift, rebuild, force, final kick per step.
    Returns (final state, last accelerations)."""
    s = np.array(state, dtype=np.float32, order="C")
    a = GpuTreeBuild(s, bucket, max_depth).accelerations(theta, softening, g)
    for _ in range(steps):
        kick_drift_f32(s, a, dt)
        a = GpuTreeBuild(s, bucket, max_depth).accelerations(theta, softening, g)
        final_kick_f32(s, a, dt)
    return s, a
