"""The serving loop: requests admitted by device-memory footprint onto concurrent streams give
the same FIDs as solo runs (and as the oracle), also when the budget forces queueing."""
import numpy as np
import pytest

from blochlib_amd import engine, serve, workloads
from blochlib_amd.powder import zcw
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_mix(n=10, seed=51):
    rng = np.random.default_rng(seed)
    pw = {1154: zcw(1154), 3722: zcw(3722)}
    out = []
    t = 0.0
    for i in range(n):
        S = int(rng.integers(2, 7))
        sy = workloads.cluster(S, rng)
        K = 1154 if i % 2 else 3722
        sub = pw[K].subset(np.arange(int(rng.integers(20, 90))) * 13 % K)
        wr = float(rng.uniform(5000, 25000))
        nst = 2 * int(rng.integers(8, 33))
        t += float(rng.exponential(0.005))
        out.append(serve.Request(i, sy, wr, workloads.MAGIC_DEG, nst * wr, int(rng.integers(64, 300)),
                                 sub, sy.op("Ix"), sy.op("Ip"), arrival=t))
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("budget_factor", [None, 1.5])
def test_mix_equals_solo_runs(budget_factor):
    reqs = small_mix()
    budget = None
    if budget_factor:
        # room for about one and a half of the largest requests: admission must queue
        fps = []
        for r in reqs:
            p = engine.Plan(r.system, r.wr, r.rotor, r.sw, r.npts)
            st = p.stats()
            fps.append(st["fixed_bytes"] + st["bytes_per_crystallite"] * min(len(r.powder), 64))
            p.close()
        budget = budget_factor * max(fps)
    print(f"[serve] {len(reqs)} requests, budget {budget}", flush=True)
    srv = serve.FIDServer(budget_bytes=budget, streams=4)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    for r in done:
        solo, sw = engine.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        assert r.sw_out == sw
        # equal to rounding: each crystallite's FID is summed in a plan-fixed grouping, but the
        # group partial sums reach it by atomic adds in run-dependent order
        err = rel(r.fid, solo)
        print(f"[serve] request {r.rid}: mix vs solo rel-L2 {err:.2e}", flush=True)
        assert err < 1e-11, r.rid
        assert r.t_done >= r.t_admit >= r.arrival - 1e-9
    # and two of them against the oracle
    for r in done[:2]:
        ref, _ = O.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        assert rel(r.fid, ref) < 1e-9
