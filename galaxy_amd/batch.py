"""Batched ensembles: many independent simulations advanced together (C ABI gbh_batch_*).

Each simulation keeps its own device buffers; one launch per phase covers all
of them and a whole step is captured in a HIP graph.  Results are bit-identical
to running each simulation alone through the drop-in API.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .bh import BhError, Config, NbodyGpu, as_state


def _check(rc):
    if rc != 0:
        raise BhError(L.lib().gbh_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Batch:
    def __init__(self, gpu: NbodyGpu, states, configs, dts, use_graph=True, sort_tile=2048,
                 traverse_block=256, traverse_waves_per_cu=32, streams=1, topo_tile=0):
        states = [as_state(s) for s in states]
        if not isinstance(configs, (list, tuple)):
            configs = [configs] * len(states)
        if not isinstance(dts, (list, tuple)):
            dts = [dts] * len(states)
        self.counts = np.array([len(s) for s in states], np.uint32)
        allst = np.ascontiguousarray(np.concatenate(states), np.float32)
        params = (L.SimParams * len(states))()
        for i, (c, dt) in enumerate(zip(configs, dts)):
            params[i].config = c.c()
            params[i].dt = float(dt)
        tuning = L.Tuning(sort_tile, traverse_block, traverse_waves_per_cu, 1 if use_graph else 0, streams,
                          topo_tile)
        h = C.c_void_p()
        _check(L.lib().gbh_batch_create(gpu._h, len(states), _p(self.counts), _p(allst), params,
                                        C.byref(tuning), C.byref(h)))
        self._h = h
        self.gpu = gpu

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def step(self, steps: int):
        _check(L.lib().gbh_batch_step(self._h, int(steps)))

    def read(self, sim: int):
        n = int(self.counts[sim])
        s = np.zeros((n, 8), np.float32)
        a = np.zeros((n, 4), np.float32)
        _check(L.lib().gbh_batch_read(self._h, sim, _p(s), _p(a)))
        return s, a

    def force_evidence(self, sim: int):
        n = int(self.counts[sim])
        st = np.zeros((n, 3), np.uint32)
        h = np.zeros(n, np.uint64)
        _check(L.lib().gbh_batch_force_evidence(self._h, sim, _p(st), _p(h)))
        return st, h

    def tree_counts(self, sim: int):
        """(cells, traversal nodes) of a simulation's current tree."""
        c, n = C.c_uint32(), C.c_uint32()
        _check(L.lib().gbh_batch_tree_counts(self._h, sim, C.byref(c), C.byref(n)))
        return int(c.value), int(n.value)

    def device_bytes(self) -> int:
        return int(L.lib().gbh_batch_device_bytes(self._h))

    def phase_ms(self):
        """Kernel time of one un-captured step per phase (advances the ensemble by one step)."""
        ms = np.zeros(8)
        k = C.c_uint32()
        _check(L.lib().gbh_batch_profile_step(self._h, _p(ms), 8, C.byref(k)))
        names = ["kick_drift_bounds", "keys", "radix_sort", "topology", "top_aggregates", "traversal"]
        return dict(zip(names, ms[: k.value].tolist()))
