"""Serving: continuous batching of 'advance this galaxy k steps' requests (C ABI gbh_server_*).

Requests are admitted FIFO while the HBM budget and slot cap allow; their
state stays resident on the device until they finish; every round advances all
in-flight requests together with one launch per phase.  A request's result is
bit-identical whether it runs alone or in a mix.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .bh import BhError, Config, NbodyGpu, as_state


def _check(rc):
    if rc != 0:
        raise BhError(L.lib().gbh_last_error().decode())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Completion:
    id: int
    submit_time: float
    admit_time: float
    finish_time: float
    n: int
    steps: int
    status: int

    @property
    def latency(self) -> float:
        return self.finish_time - self.submit_time


class Server:
    def __init__(self, gpu: NbodyGpu, hbm_budget_bytes: int = 200 * 10**9, steps_per_round: int = 10,
                 max_inflight: int = 0, sort_tile: int = 2048, traverse_block: int = 256,
                 traverse_waves_per_cu: int = 32, streams: int = 1, topo_tile: int = 0):
        opt = L.ServerOptions(int(hbm_budget_bytes), int(steps_per_round), int(max_inflight),
                              L.Tuning(sort_tile, traverse_block, traverse_waves_per_cu, 0, streams, topo_tile))
        h = C.c_void_p()
        _check(L.lib().gbh_server_create(gpu._h, C.byref(opt), C.byref(h)))
        self._h = h
        self.gpu = gpu
        self._n = {}

    @staticmethod
    def request_bytes(n: int) -> int:
        """Device bytes one resident request of n bodies holds (state + tree buffers)."""
        return int(L.lib().gbh_sim_device_bytes(int(n)))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gbh_server_destroy(self._h)  # (stops the worker thread first)
            self._h = None

    def __del__(self):
        self.close()

    def now(self) -> float:
        return float(L.lib().gbh_server_now(self._h))

    def submit(self, state, config: Config, dt: float, steps: int) -> int:
        s = as_state(state)
        p = L.SimParams(config.c(), float(dt))
        rid = C.c_uint64()
        _check(L.lib().gbh_server_submit(self._h, _p(s), len(s), C.byref(p), int(steps), C.byref(rid)))
        self._n[rid.value] = len(s)
        return int(rid.value)

    def round(self, cap: int = 4096):
        buf = (L.Completion * cap)()
        k = C.c_uint32()
        _check(L.lib().gbh_server_round(self._h, buf, cap, C.byref(k)))
        return [Completion(c.id, c.submit_time, c.admit_time, c.finish_time, c.n, c.steps, c.status)
                for c in buf[: min(k.value, cap)]]

    def start(self):
        """Continuous serving: the library's worker thread runs rounds from now on."""
        _check(L.lib().gbh_server_start(self._h))

    def stop(self):
        """Stop the worker thread after its current round (raises if it failed)."""
        _check(L.lib().gbh_server_stop(self._h))

    def poll(self, wait: float = 0.0, cap: int = 4096):
        """Completions the worker produced since the last poll; waits up to `wait`
        seconds for one when none is ready."""
        buf = (L.Completion * cap)()
        k = C.c_uint32()
        _check(L.lib().gbh_server_poll(self._h, buf, cap, C.byref(k), float(wait)))
        return [Completion(c.id, c.submit_time, c.admit_time, c.finish_time, c.n, c.steps, c.status)
                for c in buf[: k.value]]

    def stats(self):
        q, i, r = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(L.lib().gbh_server_stats(self._h, C.byref(q), C.byref(i), C.byref(r)))
        return int(q.value), int(i.value), int(r.value)

    def drain(self):
        out = []
        while True:
            q, i, _ = self.stats()
            if q == 0 and i == 0:
                return out
            out.extend(self.round())

    def take(self, rid: int, keep: bool = True):
        """Final (state, accel) of a completed request; keep=False releases it
        without copying."""
        n = self._n.pop(rid)
        if not keep:
            _check(L.lib().gbh_server_take_result(self._h, rid, None, None))
            return None
        s = np.zeros((n, 8), np.float32)
        a = np.zeros((n, 4), np.float32)
        _check(L.lib().gbh_server_take_result(self._h, rid, _p(s), _p(a)))
        return s, a
