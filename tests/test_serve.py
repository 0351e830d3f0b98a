"""Host side of the serving loop: the seeded request mix of the serving configuration."""
import numpy as np
import pytest

from blochlib_amd import serve


def test_config5_generator_shape():
    reqs = serve.config5_requests(80)
    S = np.array([r.system.spins for r in reqs])
    assert S.min() >= 2 and S.max() <= 10
    assert all(int(round(r.sw / r.wr)) % 2 == 0 for r in reqs)
    rate = len(reqs) / reqs[-1].arrival
    assert 120 < rate < 300


class _FakeDevice:
    """Streams and events of a device model on a virtual clock: each stream runs its slices in
    order, one after another, each for (crystallites x work per crystallite) / rate seconds."""

    def __init__(self, clock):
        self.clock = clock
        self.busy = {}
        self.events = {}
        self.n = 0

    def hipStreamCreateWithFlags(self, ref, flags):
        self.n += 1
        ref._obj.value = self.n
        self.busy[self.n] = 0.0
        return 0

    def hipEventCreateWithFlags(self, ref, flags):
        self.n += 1
        ref._obj.value = self.n
        self.events[self.n] = 0.0
        return 0

    def hipEventRecord(self, ev, stream):
        self.events[ev.value] = self.busy[stream.value]
        return 0

    def hipEventQuery(self, ev):
        return 0 if self.clock() >= self.events[ev.value] else 600

    def hipEventDestroy(self, ev):
        return 0

    def hipStreamDestroy(self, s):
        return 0


def _fake_plan_class(dev, log):
    class FakePlan:
        RATE = 2e9          # a 50-crystallite slice of the 8-spin request: 0.05 s

        def __init__(self, system, wr, rotor, sw, npts, precision="fp64"):
            self.S = system.spins
            self.steps = int(round(sw / wr))
            self.sw = sw
            self.npts = npts
            self.next = 0
            self.acc = 0.0

        def stats(self):
            return {"flops_per_cs": float(4 ** self.S), "fixed_bytes": 1e3,
                    "bytes_per_crystallite": 10.0, "blocks": 1}

        def set_memory_budget(self, b):
            pass

        def set_operators(self, rho, det):
            pass

        def set_powder(self, pw):
            self.K = len(pw)
            self.next = 0
            self.acc = 0.0

        def run_slice(self, c0, c1, normalize=True, stream=None):
            assert c0 == self.next and c1 > c0, "slices of one request must follow in order"
            dur = (c1 - c0) * 4 ** self.S * self.steps / self.RATE
            dev.busy[stream] = max(dev.busy[stream], dev.clock()) + dur
            self.acc += sum(range(c0, c1))
            self.next = c1
            log.append((id(self), stream))

        def fetch(self):
            assert self.next == self.K
            return np.full(self.npts, self.acc)

        def close(self):
            pass

    return FakePlan


def _fake_server(monkeypatch, **kw):
    t = [0.0]

    def clock():
        t[0] += 2e-4
        return t[0]

    dev = _FakeDevice(clock)
    log = []
    monkeypatch.setattr(serve, "_hip", lambda: dev)
    monkeypatch.setattr(serve.engine, "Plan", _fake_plan_class(dev, log))
    monkeypatch.setattr(serve.time, "sleep", lambda s: None)
    return serve.FIDServer(budget_bytes=1e9, fill_tasks=1, **kw), clock, log


def _mix():
    from blochlib_amd import workloads
    from blochlib_amd.powder import zcw
    rng = np.random.default_rng(3)
    big = workloads.cluster(8, rng)
    reqs = [serve.Request(0, big, 10000.0, workloads.MAGIC_DEG, 320000.0, 64, zcw(1154),
                          None, None, arrival=0.0)]
    for i in range(1, 5):
        sy = workloads.cluster(3, rng)
        reqs.append(serve.Request(i, sy, 10000.0, workloads.MAGIC_DEG, 320000.0, 64, zcw(1154),
                                  None, None, arrival=0.5 + 0.01 * i))
    return reqs


@pytest.mark.parametrize("policy", ["srpt", "fifo"])
def test_small_requests_overtake_a_large_one(monkeypatch, policy):
    """One stream: a large request admitted first yields its stream at a slice boundary to the
    small requests that arrive while it runs (policy srpt), and resumes where it stopped; with
    policy fifo they wait for it.  Every request's slices are issued in crystallite order."""
    srv, clock, log = _fake_server(monkeypatch, streams=1, slice_flops=4 ** 8 * 32 * 50,
                                   min_slice=1, depth=2, policy=policy)
    reqs = _mix()
    done = srv.run(reqs, clock=clock)
    by = {r.rid: r for r in done}
    assert sorted(by) == [0, 1, 2, 3, 4]
    for r in done:
        assert r.error is None
        assert np.all(r.fid == sum(range(1154)))
    big_done = by[0].t_done
    small_done = max(by[i].t_done for i in range(1, 5))
    if policy == "srpt":
        assert small_done < big_done and srv.switches >= 1
        assert max(by[i].latency for i in range(1, 5)) < 0.2 * by[0].latency
    else:
        assert small_done > big_done and srv.switches == 0


def test_server_prefers_least_remaining_work_on_free_streams(monkeypatch):
    """Two streams, queue ordered by estimated work: the four small requests share the second
    stream and finish before the large one; nothing is left in flight."""
    srv, clock, log = _fake_server(monkeypatch, streams=2, slice_flops=4 ** 8 * 32 * 50,
                                   min_slice=1)
    reqs = _mix()
    for r in reqs:
        r.arrival = 0.0
    done = srv.run(reqs, clock=clock)
    order = [r.rid for r in done]
    assert order[-1] == 0 and sorted(order) == [0, 1, 2, 3, 4]
    assert all(sl.act is None and not sl.inflight for sl in srv.slots)


def test_slice_size_keeps_per_block_launches_full(monkeypatch):
    """Slices follow slice_flops, but never leave a per-(crystallite, block) launch with fewer than
    fill_tasks workgroups, and never exceed the powder."""
    srv, clock, log = _fake_server(monkeypatch, streams=1)

    class P:
        steps = 128

        def __init__(self, per, blocks):
            self._s = {"flops_per_cs": per, "blocks": blocks}

        def stats(self):
            return self._s

    srv.fill_tasks, srv.min_slice = 1024, 4
    srv.slice_flops = 4e12
    assert srv.slice_size(P(1.5e9, 11), 3722) == 94        # fill bound: ceil(1024 / 11)
    assert srv.slice_size(P(1e6, 11), 3722) == 3722        # cheap plan: the whole powder
    assert srv.slice_size(P(1e7, 11), 3722) == 3125        # flops bound: 4e12 / 1.28e9
    assert srv.slice_size(P(1.5e9, 2000), 3722) == 20      # flops bound above fill and minimum
    assert srv.slice_size(P(1e12, 2000), 10) == 4          # at least min_slice
