"""Continuous-batching simulation server on one GPU (xo_serve_* of
libxomc.so).

A single device-owner thread holds the native server: it admits submitted
requests when their accumulators, lookup tables and voxels fit in the
server's device arenas (requests that do not fit wait, first in first out),
queues ticks of work, and copies results back as requests complete.
Client threads only enqueue requests and wait on them.

Every packet of a request draws from the generator keyed on (request seed,
packet index): a layered request's integers equal those of a batch run of
the same instance, a voxel request's those of ``mcvox.Mc.run(...,
packet_seed=seed)``, whatever else is served at the same time. The
reference has no serving layer; it runs one ``Mc.run`` per OpenCL queue at a
time (xopto/mcml/mc.py:730-1018).
"""
import collections
import ctypes
import threading
import time
from ctypes import POINTER, c_float, c_int32, c_uint32, c_uint64
from typing import Optional

import numpy as np

from . import _native as N

_LAYERED, _VOXEL = 0, 1


class Request:
    """Handle of one submitted simulation (a future)."""

    def __init__(self, kind, npackets, payload, finish):
        self.kind = kind
        self.npackets = int(npackets)
        self.id = None
        self.t_submit = time.perf_counter()
        self.t_admit = None
        self.t_done = None
        self.accum = None
        self.error = None
        self._payload = payload
        self._finish = finish
        self._result = None
        self._cancel = False
        self._event = threading.Event()

    def done(self) -> bool:
        return self._event.is_set()

    def wait(self, timeout: Optional[float] = None) -> bool:
        return self._event.wait(timeout)

    def result(self, timeout: Optional[float] = None):
        """The simulation result (Detectors for a layered request, the
        (None, fluence, detectors) tuple of Mc.run for a voxel request, the
        raw accumulator array for a pre-packed one)."""
        if not self._event.wait(timeout):
            raise TimeoutError('request not complete')
        if self.error is not None:
            raise self.error
        if self._result is None and self._finish is not None:
            self._result = self._finish(self.accum)
        return self._result if self._finish is not None else self.accum

    def cancel(self):
        self._cancel = True

    @property
    def latency(self) -> Optional[float]:
        """Seconds from submit to result (None until complete)."""
        if self.t_done is None:
            return None
        return self.t_done - self.t_submit


def _ptr(arr, ctype):
    return arr.ctypes.data_as(POINTER(ctype))


class Server:
    """Serve layered and voxelized simulations on one device.

    Capacities are in accumulator bins (u64), LUT floats and voxels; the
    remaining keyword arguments set xo_serve_opts fields (tick_packets,
    vox_tick_packets, chunk, block, use_graph, ...).
    """

    def __init__(self, device: int = 0, xinit: int = 0x5EED000000000001,
                 idle_sleep: float = 2e-4, **opts):
        self._ctx = N.Context(device, 1, xinit)
        o = N.XoServeOpts()
        N.lib().xo_serve_default_opts(ctypes.byref(o))
        for key, value in opts.items():
            if not hasattr(o, key):
                raise TypeError('unknown server option ' + key)
            setattr(o, key, value)
        self.opts = o
        h = N.lib().xo_serve_create(self._ctx._h, ctypes.byref(o))
        if not h:
            raise N.NativeError('xo_serve_create: ' + N.last_error())
        self._h = h
        self._idle_sleep = float(idle_sleep)
        self._cv = threading.Condition()
        self._incoming = collections.deque()
        self._pending = collections.deque()
        self._active = {}
        self._native_active = 0  # requests the native server still holds
        self._stop = False
        self._error = None
        self._thread = threading.Thread(target=self._loop, name='xo-serve',
                                        daemon=True)
        self._thread.start()

    # ---------------- submission (any thread) ----------------
    def _submit(self, req: Request) -> Request:
        with self._cv:
            if self._stop:
                raise RuntimeError('server is closed')
            if self._error is not None:
                raise self._error
            self._incoming.append(req)
            self._cv.notify()
        return req

    def submit_packed_layered(self, inst: N.XoBatchInstance, lut=None,
                              finish=None) -> Request:
        """A packed layered instance (detector offsets relative to
        inst.accum_base; LUT offsets into its own lut array)."""
        lut = None if lut is None or not len(lut) else \
            np.ascontiguousarray(lut, dtype=np.float32)
        return self._submit(Request(_LAYERED, inst.npackets, (inst, lut),
                                    finish))

    def submit_layered(self, instance, seed: Optional[int] = None,
                       options=None) -> Request:
        """An mcml.batch.Instance; the result is its Detectors."""
        from .mcml.batch import Batch
        if seed is not None:
            instance.seed = seed
        batch = Batch([instance], options=options)
        arr, lut, _ = batch.pack()
        inst = N.XoBatchInstance()
        ctypes.pointer(inst)[0] = arr[0]
        return self.submit_packed_layered(
            inst, lut, finish=lambda accum: batch.results(
                np.concatenate([np.zeros(arr[0].accum_base, np.uint64),
                                accum]))[0])

    def submit_packed_voxel(self, cfg: N.XoMcvoxCfg, voxels, lut,
                            naccum: int, npackets: int, seed: int,
                            finish=None) -> Request:
        voxels = np.ascontiguousarray(voxels, dtype=np.int32)
        lut = None if lut is None or not len(lut) else \
            np.ascontiguousarray(lut, dtype=np.float32)
        return self._submit(Request(
            _VOXEL, npackets,
            (cfg, voxels, lut, int(naccum), int(seed)), finish))

    def submit_voxel(self, mc, nphotons: int, seed: int) -> Request:
        """An mcvox.Mc simulation of nphotons packets keyed by seed; the
        result is Mc.run's (None, fluence, detectors)."""
        cfg, lut, naccum = mc.pack()
        nphotons = int(nphotons)
        return self.submit_packed_voxel(
            cfg, mc.voxels.material, lut, naccum, nphotons, seed,
            finish=lambda accum: mc.unpack(accum, nphotons))

    # ---------------- device-owner thread ----------------
    def _admit(self, req: Request) -> int:
        lib = N.lib()
        rid = c_uint64()
        if req.kind == _LAYERED:
            inst, lut = req._payload
            lut_p = None if lut is None else _ptr(lut, c_float)
            nlut = 0 if lut is None else lut.size
            rc = lib.xo_serve_admit_mcml(self._h, ctypes.byref(inst), lut_p,
                                         c_uint32(nlut), ctypes.byref(rid))
            naccum = int(inst.accum_count)
        else:
            cfg, voxels, lut, naccum, seed = req._payload
            lut_p = None if lut is None else _ptr(lut, c_float)
            nlut = 0 if lut is None else lut.size
            rc = lib.xo_serve_admit_mcvox(
                self._h, ctypes.byref(cfg), _ptr(voxels, c_int32),
                c_uint64(voxels.size), lut_p, c_uint32(nlut),
                c_uint64(naccum), c_uint64(req.npackets), c_uint64(seed),
                ctypes.byref(rid))
        if rc == 0:
            req.id = int(rid.value)
            req._naccum = naccum
            req.t_admit = time.perf_counter()
        elif rc < 0:
            req.error = ValueError(N.last_error())
        return rc

    def _deliver(self, rid: int):
        req = self._active.pop(rid)
        accum = np.empty(req._naccum, dtype=np.uint64)
        N.check(N.lib().xo_serve_result(self._h, c_uint64(rid),
                                        _ptr(accum, c_uint64),
                                        c_uint64(req._naccum)),
                'xo_serve_result')
        req.accum = accum
        req.t_done = time.perf_counter()
        req._payload = None
        req._event.set()

    def _loop(self):
        lib = N.lib()
        ids = (c_uint64 * 256)()
        n = c_uint32()
        active = c_uint32()
        try:
            while True:
                with self._cv:
                    while (not self._stop and not self._incoming and
                           not self._pending and not self._active and
                           not self._native_active):
                        self._cv.wait()
                    while self._incoming:
                        self._pending.append(self._incoming.popleft())
                    if self._stop and not self._active:
                        break
                # admit in arrival order; a request that does not fit
                # waits (and holds back later ones) until memory frees up
                while self._pending:
                    req = self._pending[0]
                    if req._cancel:
                        self._pending.popleft()
                        req.error = RuntimeError('cancelled')
                        req._event.set()
                        continue
                    rc = self._admit(req)
                    if rc == 1:
                        break
                    self._pending.popleft()
                    if rc < 0:
                        req._event.set()
                    else:
                        self._active[req.id] = req
                for rid, req in list(self._active.items()):
                    if req._cancel:
                        lib.xo_serve_cancel(self._h, c_uint64(rid))
                        del self._active[rid]
                        req.error = RuntimeError('cancelled')
                        req._event.set()
                N.check(lib.xo_serve_step(self._h, ctypes.byref(active)),
                        'xo_serve_step')
                self._native_active = active.value
                while True:
                    N.check(lib.xo_serve_poll(self._h, ids, 256,
                                              ctypes.byref(n)),
                            'xo_serve_poll')
                    for k in range(n.value):
                        if ids[k] in self._active:
                            self._deliver(int(ids[k]))
                    if n.value < 256:
                        break
                if self._active or self._native_active:
                    time.sleep(self._idle_sleep)
        except Exception as exc:  # device failure: fail every request
            with self._cv:
                self._error = exc
                waiting = list(self._incoming) + list(self._pending) + \
                    list(self._active.values())
                self._incoming.clear()
                self._pending.clear()
                self._active.clear()
            for req in waiting:
                req.error = exc
                req._event.set()

    # ---------------- control ----------------
    def stats(self) -> dict:
        s = N.XoServeStats()
        N.check(N.lib().xo_serve_get_stats(self._h, ctypes.byref(s)),
                'xo_serve_get_stats')
        return s.todict()

    def drain(self, timeout: Optional[float] = None) -> bool:
        """Wait until every submitted request has completed."""
        t_end = None if timeout is None else time.perf_counter() + timeout
        while True:
            with self._cv:
                busy = bool(self._incoming or self._pending or self._active
                            or self._native_active)
                if self._error is not None:
                    raise self._error
            if not busy:
                return True
            if t_end is not None and time.perf_counter() > t_end:
                return False
            time.sleep(1e-3)

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread.is_alive():
            self._thread.join()
        if getattr(self, '_h', None):
            N.lib().xo_serve_destroy(self._h)
            self._h = None
        self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_load(server: Server, make_request, qps: float, duration: float,
                 seed: int = 20261202):
    """Submit requests at Poisson arrival times for `duration` seconds.
    make_request(server, k) submits the k-th request and returns its handle.
    Returns (requests, seconds from the first submit to the last result)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t0 = time.perf_counter()
    t_next = t0
    k = 0
    while t_next - t0 < duration:
        now = time.perf_counter()
        if now < t_next:
            time.sleep(min(t_next - now, 1e-3))
            continue
        reqs.append(make_request(server, k))
        k += 1
        t_next += rng.exponential(1.0 / qps)
    for r in reqs:
        r.wait()
    t_end = max((r.t_done for r in reqs if r.t_done is not None),
                default=time.perf_counter())
    return reqs, t_end - t0


def latency_summary(reqs) -> dict:
    lat = np.array([r.latency for r in reqs if r.latency is not None])
    if lat.size == 0:
        return {'p50_ms': None, 'p99_ms': None, 'mean_ms': None, 'n': 0}
    return {'p50_ms': float(np.percentile(lat, 50) * 1e3),
            'p99_ms': float(np.percentile(lat, 99) * 1e3),
            'mean_ms': float(lat.mean() * 1e3), 'n': int(lat.size)}
