"""Serving loop: powder-FID requests admitted onto one GPU by their device-memory footprint.

Each request is one powder FID (a spin system, MAS conditions, powder, operators).  The server
keeps a FIFO of arrived requests and admits the head whenever its working storage fits the
remaining device budget and an execution stream is free; every admitted request gets a plan
sized to its share (blfid_plan_set_memory_budget, so large powders run in crystallite batches)
and runs asynchronously on its own HIP stream, so small requests overlap large ones.  A request
runs as slices of its powder (blfid_plan_run_slice) of about `slice_flops` of work each (at
least `min_slice` crystallites), issued `depth` at a time: a large request's kernels then end
every few hundred milliseconds and the device's queues interleave the kernels of requests
admitted after it, instead of those waiting for the whole large FID.  Slices give the
unsliced FID bit for bit.  Plans are cached by (system, acquisition, precision): a repeated
system reuses its device tables.  Completion is observed with HIP events; the FID is fetched and
the request's latency (arrival to result on the host) recorded.

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
    events: list                  # one HIP event per slice in flight (depth of them)
    req: Request | None = None
    plan: engine.Plan | None = None
    key: tuple | None = None
    next_c: int = 0               # first crystallite of the next slice
    slice_k: int = 0              # crystallites per slice
    inflight: list = field(default_factory=list)  # events of the slices issued, oldest first


class FIDServer:
    """Admission by device-memory footprint onto `streams` concurrent HIP streams."""

    def __init__(self, budget_bytes: float | None = None, streams: int = 4,
                 precision: str = "fp64", cache_plans: int = 4, min_batch: int = 64,
                 slice_flops: float = 4e12, min_slice: int = 64, depth: int = 2):
        self.hip = _hip()
        if budget_bytes is None:
            free, tot = ctypes.c_size_t(), ctypes.c_size_t()
            self.hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(tot))
            budget_bytes = 0.6 * free.value
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.precision = precision
        self.min_batch = min_batch
        self.slice_flops = float(slice_flops)
        self.min_slice = max(1, int(min_slice))
        self.depth = max(1, int(depth))
        self.slots = []
        for _ in range(streams):
            s = ctypes.c_void_p()
            # non-blocking streams (default priority): independent of the null stream on which
            # plan tables are uploaded
            self.hip.hipStreamCreateWithFlags(ctypes.byref(s), 1)
            evs = []
            for _ in range(self.depth):
                e = ctypes.c_void_p()
                self.hip.hipEventCreateWithFlags(ctypes.byref(e), 2)  # hipEventDisableTiming
                evs.append(e)
            self.slots.append(_Slot(s, evs))
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
            for e in sl.events:
                self.hip.hipEventDestroy(e)
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

    def slice_size(self, p: engine.Plan, K: int) -> int:
        """Crystallites per slice: about slice_flops of the plan's work, at least min_slice (so
        that every launch still spreads over the device), at most the whole powder."""
        per = max(p.stats()["flops_per_cs"] * max(p.steps, 1), 1.0)
        return int(min(max(K, 1), max(self.min_slice, self.slice_flops // per)))

    def _issue(self, sl: _Slot):
        """Issue the request's next slices while fewer than `depth` are in flight."""
        K = len(sl.req.powder)
        while sl.next_c < K and len(sl.inflight) < self.depth:
            c1 = min(K, sl.next_c + sl.slice_k)
            sl.plan.run_slice(sl.next_c, c1, True, stream=sl.stream.value)
            ev = next(e for e in sl.events if all(e.value != x.value for x in sl.inflight))
            self.hip.hipEventRecord(ev, sl.stream)
            sl.inflight.append(ev)
            sl.next_c = c1

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
                sl.req, sl.plan, sl.key = r, p, k
                sl.next_c, sl.inflight = 0, []
                try:
                    p.set_memory_budget(share)
                    p.set_operators(r.rho, r.det)
                    p.set_powder(r.powder)
                    sl.slice_k = self.slice_size(p, len(r.powder))
                    self._issue(sl)
                except engine.BlochLibError as e:
                    # the plan may hold work enqueued before the failure: it is retired, not
                    # cached, and destroyed once nothing is in flight
                    self.graveyard.append((p, 0.0))
                    self._fail(r, e, clock() - t0, done, on_done)
                    sl.req, sl.plan, sl.key, sl.inflight = None, None, None, []
                    progressed = True
                    continue
                r.footprint = share
                r.t_admit = clock() - t0
                self.used += share
                progressed = True
            for sl in self.slots:
                if sl.req is None:
                    continue
                while sl.inflight and self.hip.hipEventQuery(sl.inflight[0]) == 0:
                    sl.inflight.pop(0)
                    progressed = True
                K = len(sl.req.powder)
                if sl.next_c < K:
                    try:
                        self._issue(sl)
                    except engine.BlochLibError as e:
                        r = sl.req
                        self.used -= r.footprint
                        self.graveyard.append((sl.plan, 0.0))
                        self._fail(r, e, clock() - t0, done, on_done)
                        sl.req, sl.plan, sl.key, sl.inflight = None, None, None, []
                        progressed = True
                    continue
                if not sl.inflight:
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
