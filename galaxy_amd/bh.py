"""Python mirror of the GALAXY BH #2D backend API over the gfx950 C ABI.

Class and method names follow runtime/src/nbody_gpu.rs (NbodyGpu) and
runtime/src/nbody_parallel.rs (ParallelTreeRuntime); errors raise BhError
with the library's message (the reference's Box<dyn Error> strings).
Record arrays use the frozen transfer ABI (runtime/src/nbody_gpu.rs:22-83):

  state (n, 8) float32 = GpuEvolveState {x, y, m, 0, vx, vy, 0, 0}
  accel (n, 4) float32 = GpuAccel {ax, ay, 0, 0}
  entry (n, 4) uint32  = GpuEntry {code, body, 0, 0}
  cell  (k, 16) uint32 = GpuCell with float fields as bit patterns
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L

G_KPC3_PER_MSUN_MYR2 = 4.498502151575286e-12  # nbody/src/lib.rs:6
MAX_PARALLEL_TREE_BODIES = 65536              # nbody_parallel.rs:20


class BhError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise BhError(L.lib().gbh_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Config:
    """galaxy_nbody::Config (nbody/src/lib.rs:27-46)."""

    theta: float = 0.5
    softening_kpc: float = 0.05
    g: float = G_KPC3_PER_MSUN_MYR2
    bucket: int = 4
    max_depth: int = 48

    def c(self):
        return L.GbhConfig(self.theta, self.softening_kpc, self.g, self.bucket, self.max_depth)


def as_state(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 8:
        raise ValueError("state records must have shape (n, 8)")
    return a


class NbodyGpu:
    """Device handle (NbodyGpu::new, nbody_gpu.rs:331-621)."""

    def __init__(self, adapter: str | None = None, allow_software: bool = False):
        h = C.c_void_p()
        _check(L.lib().gbh_device_open(adapter.encode() if adapter else None, int(allow_software),
                                       C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_device_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self) -> dict:
        n = C.c_size_t()
        buf = C.create_string_buffer(2048)
        _check(L.lib().gbh_device_info_json(self._h, buf, 2048, C.byref(n)))
        return json.loads(buf.value.decode())

    # -- evolving state (nbody_gpu.rs:786-1118) --
    def create_evolving_state(self, states) -> "EvolvingState":
        return EvolvingState(self, as_state(states))

    def kick_drift(self, evolving: "EvolvingState", dt: float):
        t = L.StageTiming()
        _check(L.lib().gbh_kick_drift(evolving._h, float(dt), C.byref(t)))
        return t.as_dict()

    def final_kick(self, evolving: "EvolvingState", dt: float):
        t = L.StageTiming()
        _check(L.lib().gbh_final_kick(evolving._h, float(dt), C.byref(t)))
        return t.as_dict()

    def read_evolving_state(self, evolving: "EvolvingState") -> np.ndarray:
        out = np.zeros((evolving.count, 8), np.float32)
        _check(L.lib().gbh_state_read(evolving._h, _p(out), None))
        return out

    def read_evolving_accelerations(self, evolving: "EvolvingState") -> np.ndarray:
        out = np.zeros((evolving.count, 4), np.float32)
        _check(L.lib().gbh_accel_read(evolving._h, _p(out), None))
        return out

    # -- FP64 diagnostics (measurement; no reference counterpart) --
    def direct_f64(self, evolving: "EvolvingState", config: "Config", index) -> np.ndarray:
        """f64 direct accelerations of the listed bodies (direct_acceleration_for,
        nbody/src/lib.rs:315-329, over the state widened to f64); shape (k, 2)."""
        idx = np.ascontiguousarray(index, np.uint32)
        ax = np.zeros(len(idx), np.float64)
        ay = np.zeros(len(idx), np.float64)
        cfg = config.c()
        _check(L.lib().gbh_diag_direct(evolving._h, C.byref(cfg), len(idx), _p(idx), _p(ax), _p(ay)))
        return np.stack([ax, ay], axis=1)

    def probe_error(self, evolving: "EvolvingState", config: "Config", accel, probes: int) -> dict:
        """probe_error (nbody/src/lib.rs:370-395): RMS and max relative error of
        `accel` (rows ax, ay, ...) against f64 direct summation on probe bodies
        k*(N-1)/(probes-1)."""
        n = evolving.count
        probes = max(1, min(int(probes), n))
        den = max(probes - 1, 1)
        idx = (np.arange(probes, dtype=np.uint64) * np.uint64(n - 1)) // np.uint64(den)
        exact = self.direct_f64(evolving, config, idx.astype(np.uint32))
        a = np.asarray(accel, np.float64)[idx.astype(np.int64), :2]
        denom = np.maximum(np.hypot(exact[:, 0], exact[:, 1]), 1e-30)
        rel = np.hypot(a[:, 0] - exact[:, 0], a[:, 1] - exact[:, 1]) / denom
        return {"rms_relative": float(np.sqrt(np.mean(rel * rel))), "max_relative": float(rel.max()),
                "probes": probes}

    def energy(self, evolving: "EvolvingState", softening: float, g: float) -> dict:
        """Total energy (kinetic + Plummer-softened pair potential), Lz and kinetic
        energy of the resident state, f64."""
        out = np.zeros(3, np.float64)
        _check(L.lib().gbh_diag_energy(evolving._h, float(softening), float(g), _p(out)))
        return {"energy": float(out[0]), "lz": float(out[1]), "kinetic": float(out[2])}

    # -- BH #2B1 / #2B2 host-tree ingest (nbody_gpu.rs:640-938) --
    # -- BH #2C GPU tree (nbody_gpu.rs:1145-1431) --
    def create_gpu_tree(self, evolving: "EvolvingState") -> "GpuTree":
        return GpuTree(self, evolving)

    def rebuild_gpu_tree(self, evolving: "EvolvingState", tree: "GpuTree", config: "Config") -> dict:
        t = L.BuildTiming()
        _check(L.lib().gbh_gpu_tree_rebuild(evolving._h, tree._h, C.byref(config.c()), C.byref(t)))
        return t.as_dict()

    def force_from_gpu_tree(self, evolving: "EvolvingState", tree: "GpuTree", config: "Config") -> dict:
        t = L.StageTiming()
        _check(L.lib().gbh_gpu_tree_force(evolving._h, tree._h, C.byref(config.c()), C.byref(t)))
        return t.as_dict()

    def read_gpu_tree_evidence(self, tree: "GpuTree"):
        """GpuTreeEvidence (nbody_gpu.rs:119-123) -> (meta u32[8], entries u32[n, 4],
        cells u32[k, 16] in level order)."""
        meta = np.zeros(8, np.uint32)
        cap = 16 * tree.count + 1
        entries = np.zeros((tree.count, 4), np.uint32)
        cells = np.zeros((cap, 16), np.uint32)
        k = C.c_uint32()
        _check(L.lib().gbh_gpu_tree_read_evidence(tree._h, _p(meta), _p(entries), _p(cells), cap, C.byref(k)))
        return meta, entries, cells[: k.value]

    def gpu_tree_buffer_bytes(self, tree: "GpuTree") -> int:
        return int(L.lib().gbh_gpu_tree_buffer_bytes(tree._h))

    def traverse(self, settings, bodies, entries, cells) -> np.ndarray:
        settings = np.ascontiguousarray(settings, np.uint32)
        bodies = np.ascontiguousarray(bodies, np.uint32)
        entries = np.ascontiguousarray(entries, np.uint32)
        cells = np.ascontiguousarray(cells, np.uint32)
        out = np.zeros((int(settings[0]), 4), np.float32)
        _check(L.lib().gbh_traverse_packed(self._h, _p(settings), _p(bodies), _p(entries), _p(cells),
                                           _p(out), None))
        return out

    def force_into(self, evolving: "EvolvingState", settings, bodies, entries, cells):
        settings = np.ascontiguousarray(settings, np.uint32)
        bodies = np.ascontiguousarray(bodies, np.uint32)
        entries = np.ascontiguousarray(entries, np.uint32)
        cells = np.ascontiguousarray(cells, np.uint32)
        _check(L.lib().gbh_force_packed(evolving._h, _p(settings), _p(bodies), _p(entries), _p(cells),
                                        None))


class GpuTree:
    """BH #2C correctness-first GPU tree (GpuTree, nbody_gpu.rs:107-116)."""

    def __init__(self, gpu: NbodyGpu, evolving: "EvolvingState"):
        h = C.c_void_p()
        _check(L.lib().gbh_gpu_tree_create(gpu._h, evolving._h, C.byref(h)))
        self._h = h
        self.count = evolving.count
        self._keep = evolving

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_gpu_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class EvolvingState:
    """EvolvingState (nbody_gpu.rs:132-138)."""

    def __init__(self, gpu: NbodyGpu, states: np.ndarray):
        self.gpu = gpu
        self.count = len(states)
        h = C.c_void_p()
        _check(L.lib().gbh_state_create(gpu._h, _p(states), self.count, C.byref(h), None))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_state_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def buffer_bytes(self) -> int:
        return int(L.lib().gbh_state_bytes(self._h))


@dataclass
class TreeEvidence:
    """ParallelGpuTreeEvidence (nbody_parallel.rs:83-91)."""

    meta: np.ndarray       # (8,) uint32: data[4] + bounds bits[4]
    entries: np.ndarray    # (n, 4) uint32
    cells: np.ndarray      # (active, 16) uint32
    active_cell_count: int
    leaf_count: int
    max_depth: int

    @property
    def bounds(self) -> np.ndarray:
        return self.meta[4:8].view(np.float32)


class ParallelGpuTree:
    def __init__(self, runtime: "ParallelTreeRuntime", evolving: EvolvingState, large: bool = False):
        self.runtime = runtime
        self.evolving = evolving
        self.count = evolving.count
        h = C.c_void_p()
        fn = L.lib().gbh_tree_create_large if large else L.lib().gbh_tree_create
        _check(fn(runtime._h, evolving._h, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class ParallelTreeRuntime:
    """BH #2D parallel tree runtime (nbody_parallel.rs:93-993)."""

    def __init__(self, gpu: NbodyGpu):
        self.gpu = gpu
        h = C.c_void_p()
        _check(L.lib().gbh_runtime_create(gpu._h, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_runtime_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def create_tree(self, gpu: NbodyGpu, evolving: EvolvingState, large: bool = False) -> ParallelGpuTree:
        return ParallelGpuTree(self, evolving, large)

    def rebuild(self, gpu, evolving: EvolvingState, tree: ParallelGpuTree, config: Config) -> dict:
        t = L.BuildTiming()
        _check(L.lib().gbh_tree_rebuild(self._h, evolving._h, tree._h, C.byref(config.c()), C.byref(t)))
        return t.as_dict()

    def force(self, gpu, evolving: EvolvingState, tree: ParallelGpuTree, config: Config) -> dict:
        t = L.StageTiming()
        _check(L.lib().gbh_force(self._h, evolving._h, tree._h, C.byref(config.c()), C.byref(t)))
        return t.as_dict()

    def force_evidence(self, evolving: EvolvingState, tree: ParallelGpuTree, config: Config):
        """Per-target (visited, accepted, direct) counts and interaction-list hashes."""
        stats = np.zeros((evolving.count, 3), np.uint32)
        hsh = np.zeros(evolving.count, np.uint64)
        _check(L.lib().gbh_force_evidence(self._h, evolving._h, tree._h, C.byref(config.c()),
                                          _p(stats), _p(hsh)))
        return stats, hsh

    def force_evidence_noalloc(self, evolving: EvolvingState, tree: ParallelGpuTree, config: Config):
        """force_evidence through the entry point that allocates no device memory
        (evidence staged in the library's per-GPU scratch array; at most 65,536 bodies)."""
        stats = np.zeros((evolving.count, 3), np.uint32)
        hsh = np.zeros(evolving.count, np.uint64)
        _check(L.lib().gbh_force_evidence_noalloc(self._h, evolving._h, tree._h, C.byref(config.c()),
                                                  _p(stats), _p(hsh)))
        return stats, hsh

    def counts(self, tree: ParallelGpuTree):
        a, lf, md = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(L.lib().gbh_tree_counts(tree._h, C.byref(a), C.byref(lf), C.byref(md)))
        return int(a.value), int(lf.value), int(md.value)

    def read_evidence(self, gpu, tree: ParallelGpuTree) -> TreeEvidence:
        active, _, _ = self.counts(tree)
        meta = np.zeros(8, np.uint32)
        entries = np.zeros((tree.count, 4), np.uint32)
        cells = np.zeros((active, 16), np.uint32)
        a, lf, md = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(L.lib().gbh_tree_read_evidence(tree._h, _p(meta), _p(entries), _p(cells), active,
                                              C.byref(a), C.byref(lf), C.byref(md)))
        return TreeEvidence(meta, entries, cells, int(a.value), int(lf.value), int(md.value))

    def read_flat(self, tree: ParallelGpuTree):
        """Tree in BH #2A preorder packed like pack_flat_tree -> (entries, cells, position_by_body)."""
        active, _, _ = self.counts(tree)
        entries = np.zeros((tree.count, 4), np.uint32)
        cells = np.zeros((active, 16), np.uint32)
        pos = np.zeros(tree.count, np.uint32)
        k = C.c_uint32()
        _check(L.lib().gbh_tree_read_flat(tree._h, _p(entries), _p(cells), active, C.byref(k), _p(pos)))
        return entries, cells[: k.value], pos

    def last_force_ms(self, tree: ParallelGpuTree) -> float:
        """GPU time of the last force traversal kernel (HIP events)."""
        return float(L.lib().gbh_tree_last_force_ms(tree._h))

    def buffer_bytes(self, tree: ParallelGpuTree) -> int:
        return int(L.lib().gbh_tree_buffer_bytes(tree._h))


def host_flat_tree(x, y, mass, config: Config):
    """build_flat_tree + pack_flat_tree on the host (nbody/src/flat.rs:127-268,
    nbody_gpu.rs:175-250) -> (settings u32[8], bodies u32[n, 8], entries u32[n, 4],
    cells u32[k, 16]), the records gbh_force_packed takes."""
    x, y, mass = (np.ascontiguousarray(v, np.float64) for v in (x, y, mass))
    n = len(x)
    cap = 16 * n + 1
    settings = np.zeros(8, np.uint32)
    bodies = np.zeros((n, 8), np.uint32)
    entries = np.zeros((n, 4), np.uint32)
    cells = np.zeros((cap, 16), np.uint32)
    k = C.c_uint32()
    _check(L.lib().gbh_host_flat_tree(n, _p(x), _p(y), _p(mass), C.byref(config.c()), _p(settings), _p(bodies),
                                      _p(entries), _p(cells), cap, C.byref(k)))
    return settings, bodies, entries, cells[: k.value]


def sparse_layout_bytes(n: int) -> int:
    return int(L.lib().gbh_sparse_layout_bytes(n))
