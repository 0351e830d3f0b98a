"""Serving loop: powder-FID requests admitted onto one GPU by their device-memory footprint.

Each request is one powder FID (a spin system, MAS conditions, powder, operators).  Arrived
requests wait in a queue ordered by an estimate of their work; the server admits them while
their working storage fits the remaining device budget (a plan sized to one slice of the
request, blfid_plan_set_memory_budget) and keeps every admitted request resident.  A request
runs as slices of its powder (blfid_plan_run_slice) of about `slice_flops` of work each, issued
at most `depth` at a time on one of `streams` HIP streams.  Streams go to the admitted requests
with the least remaining work: at every slice boundary a running request whose remaining work is
larger than that of a request waiting for a stream stops issuing, drains its slices in flight
and yields the stream (its plan, group accumulators included, stays resident, and it resumes
where it stopped, on whichever stream comes free).  Small requests that arrive while a 10-spin
FID is running therefore finish within a few slices instead of behind the whole large FID.
Slices give the unsliced FID bit for bit, whatever the interleaving.  Plans are cached by
(system, acquisition, precision): a repeated system reuses its device tables.  Completion is
observed with HIP events; the FID is fetched and the request's latency (arrival to result on
the host) recorded.

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

    def estimate(self) -> float:
        """Work estimate used to order the queue before a plan exists: per crystallite-step,
        the cube of every magnetisation sector of the spins (the blocks of a homonuclear
        system), times crystallites and steps per period."""
        from math import comb
        S = self.system.spins
        n = max(1, int(round(self.sw / self.wr))) if self.wr else 1
        return float(sum(comb(S, k) ** 3 for k in range(S + 1))) * len(self.powder) * n


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
class _Active:
    """An admitted request: its plan stays resident until the last slice completes."""
    req: Request
    plan: engine.Plan
    key: tuple
    share: float                  # device bytes counted against the budget
    per_c: float                  # work per crystallite (flops, all steps of a period)
    slice_k: int                  # crystallites per slice
    next_c: int = 0               # first crystallite of the next slice
    slot: "_Slot | None" = None   # the stream it runs on, or None while it waits

    @property
    def K(self) -> int:
        return len(self.req.powder)

    @property
    def remaining(self) -> float:
        return (self.K - self.next_c) * self.per_c


@dataclass
class _Slot:
    stream: ctypes.c_void_p
    events: list                  # one HIP event per slice in flight (depth of them)
    act: _Active | None = None
    inflight: list = field(default_factory=list)  # events of the slices issued, oldest first


class FIDServer:
    """Admission by device-memory footprint; streams go to the admitted requests with the least
    remaining work, switched at slice boundaries."""

    def __init__(self, budget_bytes: float | None = None, streams: int = 4,
                 precision: str = "fp64", cache_plans: int = 4, max_batch: int = 256,
                 slice_flops: float = 4e12, min_slice: int = 4, fill_tasks: int = 1024,
                 depth: int = 2, max_active: int = 32, policy: str = "srpt"):
        if policy not in ("srpt", "fifo"):
            raise ValueError("policy is 'srpt' (least remaining work first) or 'fifo'")
        self.hip = _hip()
        if budget_bytes is None:
            free, tot = ctypes.c_size_t(), ctypes.c_size_t()
            self.hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(tot))
            budget_bytes = 0.6 * free.value
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.precision = precision
        self.max_batch = max(1, int(max_batch))
        self.slice_flops = float(slice_flops)
        self.min_slice = max(1, int(min_slice))
        self.fill_tasks = max(1, int(fill_tasks))
        self.depth = max(1, int(depth))
        self.max_active = max(1, int(max_active))
        self.policy = policy
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
        self.switches = 0             # times a running request yielded its stream

    def close(self):
        for sl in self.slots:
            if sl.act is not None:
                sl.act.plan.fetch()
                sl.act.plan.close()
                sl.act = None
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
        if k in self.cache:               # a twin request's plan is cached already
            self.graveyard.append((p, nbytes))
            self.used += nbytes
        else:
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

    def slice_size(self, p: engine.Plan, K: int) -> int:
        """Crystallites per slice: about slice_flops of the plan's work, but enough that the
        per-(crystallite, block) launches (the eigen-solver's, one workgroup each) still have
        fill_tasks workgroups, at least min_slice, at most the whole powder."""
        st = p.stats()
        per = max(st["flops_per_cs"] * max(p.steps, 1), 1.0)
        fill = -(-self.fill_tasks // max(int(st["blocks"]), 1))
        return int(min(max(K, 1), max(self.min_slice, fill, self.slice_flops // per)))

    def footprint(self, p: engine.Plan, K: int, slice_k: int | None = None) -> float:
        """Device bytes to admit a request: fixed scratch + one batch of crystallites (a slice,
        at most max_batch crystallites, at most the whole powder)."""
        st = p.stats()
        kb = min(K, self.max_batch, slice_k if slice_k else K)
        return st["fixed_bytes"] + st["bytes_per_crystallite"] * max(kb, 1)

    def _issue(self, sl: _Slot, rival: float):
        """Issue the bound request's next slices while fewer than `depth` are in flight and no
        waiting request has less remaining work than it (`rival`)."""
        a = sl.act
        while a.next_c < a.K and len(sl.inflight) < self.depth:
            if self.policy == "srpt" and rival < a.remaining:
                return
            c1 = min(a.K, a.next_c + a.slice_k)
            a.plan.run_slice(a.next_c, c1, True, stream=sl.stream.value)
            ev = next(e for e in sl.events if all(e.value != x.value for x in sl.inflight))
            self.hip.hipEventRecord(ev, sl.stream)
            sl.inflight.append(ev)
            a.next_c = c1

    def _fail(self, r: Request, err: Exception, t: float, done: list, on_done):
        """Completes a request with its error (bad system, unsupported input, out of memory):
        the server and the other requests go on."""
        r.error = err
        r.fid = None
        r.t_done = t
        done.append(r)
        if on_done:
            on_done(r)

    def _order(self, r: Request):
        return (r.estimate(), r.arrival) if self.policy == "srpt" else (r.arrival,)

    def _admit(self, queue: list, active: list, now, done, on_done) -> bool:
        """Admit queued requests (least estimated work first) while their footprint fits the
        budget; returns whether anything changed."""
        progressed = False
        queue.sort(key=self._order)
        while queue and len(active) < self.max_active:
            r = queue[0]
            try:
                k, p, held = self._plan(r)
                slice_k = self.slice_size(p, len(r.powder))
                need = self.footprint(p, len(r.powder), slice_k)
            except engine.BlochLibError as e:
                queue.pop(0)
                self._fail(r, e, now(), done, on_done)
                progressed = True
                continue
            while need > self.budget - self.used and self._evict():
                pass
            if need > self.budget - self.used and not any(x.act for x in self.slots):
                self._bury()
            if need > self.budget - self.used and (active or self.used > 0):
                self._release(k, p, held)   # waits for memory; the arena goes back to the cache
                break
            queue.pop(0)
            try:
                p.set_memory_budget(need)
                p.set_operators(r.rho, r.det)
                p.set_powder(r.powder)
            except engine.BlochLibError as e:
                # the plan may hold work enqueued before the failure: it is retired, not
                # cached, and destroyed once nothing is in flight
                self.graveyard.append((p, 0.0))
                self._fail(r, e, now(), done, on_done)
                progressed = True
                continue
            per = max(p.stats()["flops_per_cs"] * max(p.steps, 1), 1.0)
            r.footprint = need
            r.t_admit = now()
            self.used += need
            active.append(_Active(r, p, k, need, per, slice_k))
            progressed = True
        return progressed

    def _retire(self, a: _Active, err, now, done, on_done):
        """Take a request off the device: completed (err None) or failed."""
        r = a.req
        self.used -= a.share
        if a.slot is not None:
            a.slot.act = None
            a.slot.inflight = []
            a.slot = None
        if err is None:
            try:
                r.fid = a.plan.fetch()
                r.sw_out = a.plan.sw
                r.t_done = now()
                self._release(a.key, a.plan, a.share)
                done.append(r)
                if on_done:
                    on_done(r)
                return
            except engine.BlochLibError as e:
                err = e
        self.graveyard.append((a.plan, 0.0))
        self._fail(r, err, now(), done, on_done)

    def run(self, requests, clock=time.perf_counter, on_done=None) -> list:
        """Serve `requests` (arrival times in seconds from now); returns them completed (a
        request that failed has `error` set and no FID)."""
        todo = deque(sorted(requests, key=lambda r: r.arrival))
        queue: list = []
        active: list = []
        done = []
        t0 = clock()

        def now():
            return clock() - t0

        while todo or queue or active:
            t = now()
            while todo and todo[0].arrival <= t:
                queue.append(todo.popleft())
            progressed = self._admit(queue, active, now, done, on_done) if queue else False
            # slices that completed; requests whose last slice completed
            for sl in self.slots:
                while sl.inflight and self.hip.hipEventQuery(sl.inflight[0]) == 0:
                    sl.inflight.pop(0)
                    progressed = True
                a = sl.act
                if a is not None and not sl.inflight and a.next_c >= a.K:
                    active.remove(a)
                    self._retire(a, None, now, done, on_done)
                    progressed = True
            waiting = sorted((a for a in active if a.slot is None),
                             key=(lambda a: (a.remaining, a.req.arrival)) if self.policy == "srpt"
                             else (lambda a: a.req.arrival))
            # a stream whose request drained at a slice boundary goes to a waiting request with
            # less remaining work
            if self.policy == "srpt":
                for sl in self.slots:
                    a = sl.act
                    if a is None or sl.inflight or not waiting:
                        continue
                    if waiting[0].remaining < a.remaining:
                        a.slot, sl.act = None, None
                        waiting.append(a)
                        waiting.sort(key=lambda x: (x.remaining, x.req.arrival))
                        self.switches += 1
                        progressed = True
            for sl in self.slots:
                if sl.act is None and waiting:
                    a = waiting.pop(0)
                    sl.act, a.slot, sl.inflight = a, sl, []
                    progressed = True
            rival = waiting[0].remaining if waiting else float("inf")
            for sl in self.slots:
                a = sl.act
                if a is None:
                    continue
                before = a.next_c
                try:
                    self._issue(sl, rival)
                except engine.BlochLibError as e:
                    active.remove(a)
                    self._retire(a, e, now, done, on_done)
                    progressed = True
                    continue
                progressed = progressed or a.next_c != before
            if self.graveyard and not any(sl.act is not None for sl in self.slots):
                self._bury()
            if not progressed:
                if queue or active:
                    time.sleep(0.0002)
                elif todo:
                    time.sleep(max(0.0, min(0.01, todo[0].arrival - now())))
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
