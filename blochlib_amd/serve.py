"""Serving loop: powder-FID requests admitted onto one GPU by their device-memory footprint.

Each request is one powder FID (a spin system, MAS conditions, powder, operators).  The server
keeps a FIFO of arrived requests and admits the head whenever its working storage fits the
remaining device budget and an execution stream is free; every admitted request gets a plan
sized to its share (blfid_plan_set_memory_budget, so large powders run in crystallite batches)
and runs asynchronously on its own HIP stream, so small requests overlap large ones.  Plans are
cached by (system, acquisition, precision): a repeated system reuses its device tables.
Completion is observed with HIP events; the FID is fetched and the request's latency (arrival
to result on the host) recorded.

This replaces nothing in BlochLib (which computes one simulation per process); it is the
service shape of the batched engine: FIDs/s and latency percentiles under an arrival process.
"""
from __future__ import annotations

import ctypes
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import engine
from .powder import Powder
from .spinsys import SolidSys


@dataclass
class Request:
    rid: int
    system: SolidSys
    wr: float
    rotor: float
    sw: float
    npts: int
    powder: Powder
    rho: np.ndarray
    det: np.ndarray
    arrival: float = 0.0           # seconds after the start of the run
    # filled by the server
    fid: np.ndarray | None = None
    sw_out: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    footprint: float = 0.0
    error: Exception | None = None  # a failed request carries its error; the others go on

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def _hip():
    L = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
    L.hipStreamDestroy.argtypes = [vp]
    L.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
    L.hipEventRecord.argtypes = [vp, vp]
    L.hipEventQuery.argtypes = [vp]
    L.hipEventDestroy.argtypes = [vp]
    L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return L


@dataclass
class _Slot:
    stream: ctypes.c_void_p
    event: ctypes.c_void_p
    req: Request | None = None
    plan: engine.Plan | None = None
    key: tuple | None = None


class FIDServer:
    """Admission by device-memory footprint onto `streams` concurrent HIP streams."""

    def __init__(self, budget_bytes: float | None = None, streams: int = 4,
                 precision: str = "fp64", cache_plans: int = 4, min_batch: int = 64):
        self.hip = _hip()
        if budget_bytes is None:
            free, tot = ctypes.c_size_t(), ctypes.c_size_t()
            self.hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(tot))
            budget_bytes = 0.6 * free.value
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.precision = precision
        self.min_batch = min_batch
        self.slots = []
        for _ in range(streams):
            s, e = ctypes.c_void_p(), ctypes.c_void_p()
            # non-blocking streams (default priority): independent of the null stream on which
            # plan tables are uploaded
            self.hip.hipStreamCreateWithFlags(ctypes.byref(s), 1)
            self.hip.hipEventCreateWithFlags(ctypes.byref(e), 2)  # hipEventDisableTiming
            self.slots.append(_Slot(s, e))
        self.cache: OrderedDict = OrderedDict()
        self.cache_max = cache_plans
        # evicted plans: freeing device memory synchronises the device, so they are destroyed
        # only while nothing is in flight (their bytes stay counted until then)
        self.graveyard = []

    def close(self):
        for sl in self.slots:
            if sl.plan is not None:
                sl.plan.fetch()
                sl.plan.close()
        for p, _ in self.cache.values():
            p.close()
        self.cache.clear()
        self._bury()
        for sl in self.slots:
            self.hip.hipEventDestroy(sl.event)
            self.hip.hipStreamDestroy(sl.stream)
        self.slots = []

    @staticmethod
    def key(r: Request, precision: str) -> tuple:
        return (r.system.to_text() if hasattr(r.system, "to_text") else id(r.system), r.wr,
                r.rotor, r.sw, r.npts, precision)

    def _plan(self, r: Request):
        """A plan for the request and the bytes it held in the cache: a cached one (its arena
        bytes leave the cache account, and go back with it if it is not admitted after all) or
        a new one (0 bytes)."""
        k = self.key(r, self.precision)
        hit = self.cache.pop(k, None)
        if hit is not None:
            self.used -= hit[1]
            return k, hit[0], hit[1]
        return k, engine.Plan(r.system, r.wr, r.rotor, r.sw, r.npts, precision=self.precision), 0.0

    def _release(self, k, p, nbytes):
        """Keep a finished plan (its device arena stays allocated and counted) or drop it."""
        self.cache[k] = (p, nbytes)
        self.used += nbytes
        while len(self.cache) > self.cache_max:
            self._evict()

    def _evict(self) -> bool:
        if not self.cache:
            return False
        _, (old, nb) = self.cache.popitem(last=False)
        self.graveyard.append((old, nb))
        return True

    def _bury(self):
        for p, nb in self.graveyard:
            p.close()
            self.used -= nb
        self.graveyard = []

    def footprint(self, p: engine.Plan, K: int) -> float:
        """Device bytes to admit a request: fixed scratch + a batch of crystallites (at least
        min_batch or the whole powder, whichever is smaller)."""
        st = p.stats()
        return st["fixed_bytes"] + st["bytes_per_crystallite"] * min(K, self.min_batch)

    def _fail(self, r: Request, err: Exception, t: float, done: list, on_done):
        """Completes a request with its error (bad system, unsupported input, out of memory):
        the server and the other requests go on."""
        r.error = err
        r.fid = None
        r.t_done = t
        done.append(r)
        if on_done:
            on_done(r)

    def run(self, requests, clock=time.perf_counter, on_done=None) -> list:
        """Serve `requests` (arrival times in seconds from now); returns them completed (a
        request that failed has `error` set and no FID)."""
        todo = deque(sorted(requests, key=lambda r: r.arrival))
        queue: deque = deque()
        done = []
        t0 = clock()
        while todo or queue or any(sl.req is not None for sl in self.slots):
            now = clock() - t0
            while todo and todo[0].arrival <= now:
                queue.append(todo.popleft())
            progressed = False
            # admit the queue head while it fits
            while queue:
                free = [sl for sl in self.slots if sl.req is None]
                if not free:
                    break
                r = queue[0]
                try:
                    k, p, held = self._plan(r)
                    need = self.footprint(p, len(r.powder))
                except engine.BlochLibError as e:
                    queue.popleft()
                    self._fail(r, e, clock() - t0, done, on_done)
                    progressed = True
                    continue
                while need > self.budget - self.used and self._evict():
                    pass
                if need > self.budget - self.used and not any(x.req for x in self.slots):
                    self._bury()
                room = self.budget - self.used
                if need > room and self.used > 0:
                    self._release(k, p, held)
                    break
                share = max(need, min(room, need * 4))
                queue.popleft()
                sl = free[0]
                try:
                    p.set_memory_budget(share)
                    p.set_operators(r.rho, r.det)
                    p.set_powder(r.powder)
                    p.run(True, stream=sl.stream.value)
                except engine.BlochLibError as e:
                    # the plan may hold work enqueued before the failure: it is retired, not
                    # cached, and destroyed once nothing is in flight
                    self.graveyard.append((p, 0.0))
                    self._fail(r, e, clock() - t0, done, on_done)
                    progressed = True
                    continue
                self.hip.hipEventRecord(sl.event, sl.stream)
                sl.req, sl.plan, sl.key = r, p, k
                r.footprint = share
                r.t_admit = clock() - t0
                self.used += share
                progressed = True
            for sl in self.slots:
                if sl.req is not None and self.hip.hipEventQuery(sl.event) == 0:
                    r = sl.req
                    self.used -= r.footprint
                    try:
                        r.fid = sl.plan.fetch()
                        r.sw_out = sl.plan.sw
                        r.t_done = clock() - t0
                        self._release(sl.key, sl.plan, r.footprint)
                        done.append(r)
                        if on_done:
                            on_done(r)
                    except engine.BlochLibError as e:
                        self.graveyard.append((sl.plan, 0.0))
                        self._fail(r, e, clock() - t0, done, on_done)
                    sl.req, sl.plan, sl.key = None, None, None
                    progressed = True
            if self.graveyard and not any(sl.req is not None for sl in self.slots):
                self._bury()
            if not progressed:
                if queue or any(sl.req is not None for sl in self.slots):
                    time.sleep(0.0002)
                elif todo:
                    time.sleep(max(0.0, min(0.01, todo[0].arrival - (clock() - t0))))
        return done


def config5_requests(n: int, rate: float = 200.0, seed: int = 1000 * 5, max_spins: int = 10):
    """The serving mix: Poisson arrivals at `rate`/s; per request a seeded 1H cluster of S spins
    (S log-uniform over 2..max_spins), MAS wr ~ U(5, 25) kHz at the magic angle, ZCW 1154 or
    3722, n ~ 2 U{16..64} steps per period (even), P ~ U[256, 2048] points; rho = Ix,
    det = I+."""
    from . import workloads
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    pows = {}
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        S = int(np.clip(np.floor(np.exp(rng.uniform(np.log(2), np.log(max_spins + 1)))), 2,
                        max_spins))
        wr = float(rng.uniform(5000.0, 25000.0))
        K = 1154 if rng.uniform() < 0.5 else 3722
        nst = 2 * int(rng.integers(16, 65))
        P = int(rng.integers(256, 2049))
        sy = workloads.cluster(S, rng)
        if K not in pows:
            pows[K] = workloads.zcw(K)
        out.append(Request(i, sy, wr, workloads.MAGIC_DEG, nst * wr, P, pows[K], sy.op("Ix"),
                           sy.op("Ip"), arrival=t))
    return out
