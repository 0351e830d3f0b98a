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


@pytest.mark.parametrize("budget_factor,sliced", [(None, False), (1.5, False), (None, True)])
def test_mix_equals_solo_runs(budget_factor, sliced):
    """sliced: every request cut into slices of 7 crystallites (blfid_plan_run_slice), two in
    flight per stream, so the slices of the four streams interleave on the device."""
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
    kw = dict(slice_flops=1.0, min_slice=7) if sliced else {}
    srv = serve.FIDServer(budget_bytes=budget, streams=4, **kw)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    for r in done:
        solo, sw = engine.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        assert r.sw_out == sw
        # bit for bit: every partial FID has one writer and the powder sum runs in a fixed
        # order, whatever the stream, the batch size or the work running beside it
        assert r.error is None, (r.rid, r.error)
        print(f"[serve] request {r.rid}: mix vs solo rel-L2 {rel(r.fid, solo):.2e}", flush=True)
        assert np.array_equal(r.fid, solo), r.rid
        assert r.t_done >= r.t_admit >= r.arrival - 1e-9
    # and two of them against the oracle
    for r in done[:2]:
        ref, _ = O.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        assert rel(r.fid, ref) < 1e-9


def test_bad_requests_fail_alone():
    """A request rejected when its plan is made (eta > 1) and one rejected when it is run
    (Tr(det det^H) = 0 with normalisation) complete with their errors; the rest of the mix is
    served and equals its solo runs."""
    from blochlib_amd.spinsys import SolidSys
    reqs = small_mix(n=6, seed=77)
    bad = SolidSys(isotopes=["1H"] * 2, csa=[(0.0, 100.0, 1.5, 0, 0.0, 0.0, 0.0)])
    r = reqs[1]
    reqs[1] = serve.Request(1, bad, r.wr, r.rotor, r.sw, r.npts, r.powder,
                            np.eye(4, dtype=complex), np.eye(4, dtype=complex), arrival=r.arrival)
    r = reqs[3]
    reqs[3] = serve.Request(3, r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho,
                            np.zeros_like(r.det), arrival=r.arrival)
    print(f"[serve] {len(reqs)} requests, two bad", flush=True)
    srv = serve.FIDServer(streams=4)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert sorted(x.rid for x in done) == list(range(len(reqs)))
    by = {x.rid: x for x in done}
    assert isinstance(by[1].error, engine.EngineError) and by[1].error.status == 1
    assert isinstance(by[3].error, engine.EngineError) and by[3].error.status == 1
    for i in (0, 2, 4, 5):
        x = by[i]
        assert x.error is None, (i, x.error)
        solo, _ = engine.powder_fid(x.system, x.wr, x.rotor, x.sw, x.npts, x.powder, x.rho, x.det)
        assert np.array_equal(x.fid, solo), i


@pytest.mark.parametrize("precision", ["fp64", "c64"])
def test_config5_stream_prefix(precision):
    """The serving configuration's own seeded stream (Poisson 200/s, 2-10-spin clusters,
    ZCW 1154/3722, 32-128 steps, 256-2048 points): its first six requests, which include a
    9-spin one (N = 512, blocks of 126 states, ZCW 3722), served with their arrival times.  Every
    request equals its solo run bit for bit, and the two smallest match the oracle (1e-9 in FP64,
    1e-5 in complex64)."""
    reqs = serve.config5_requests(6)
    assert max(r.system.spins for r in reqs) >= 9
    print(f"[serve] config-5 prefix {precision}: spins {[r.system.spins for r in reqs]}", flush=True)
    srv = serve.FIDServer(streams=4, precision=precision)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    for r in done:
        assert r.error is None, (r.rid, r.error)
        solo, sw = engine.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det,
                                     precision=precision)
        print(f"[serve] request {r.rid} S={r.system.spins}: mix == solo "
              f"{np.array_equal(r.fid, solo)}", flush=True)
        assert r.sw_out == sw
        assert np.array_equal(r.fid, solo), r.rid
    tol = 1e-9 if precision == "fp64" else 1e-5
    for r in sorted(done, key=lambda x: x.system.N * len(x.powder))[:2]:
        ref, _ = O.powder_fid(r.system, r.wr, r.rotor, r.sw, r.npts, r.powder, r.rho, r.det)
        err = rel(r.fid, ref)
        print(f"[serve] request {r.rid} S={r.system.spins}: rel-L2 vs oracle {err:.2e}", flush=True)
        assert err < tol
