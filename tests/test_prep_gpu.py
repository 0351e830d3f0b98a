"""Pulses before a plain gamma-COMPUTE fid() (SequenceRun::doFID, sequencerun_funcs.cc:693-703):
every crystallite's rho is propagated through the pulses (the engine's preparation stage,
blfid_plan_set_prep: FP64 pulse propagators, the prepared rho formed on the FID's block pairs)
and its FID runs from the pulses' end time.  GPU through the C ABI against the oracle's
orc_prep_fid (itself pinned to an explicit-propagation restatement in test_oracle.py): FP64 to
1e-9, complex64 to 1e-5."""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import zcw
from blochlib_amd.spinsys import SolidSys
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MAGIC = workloads.MAGIC_DEG


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run(name, sy, wr, sw, npts, pw, rho, det, dts, rf, precision, tol, t1=None):
    print(f"[prep] {name} {precision}: S={sy.spins} K={len(pw)} steps={len(dts)}", flush=True)
    t0 = float(np.sum(dts)) if t1 is None else float(t1[-1] + dts[-1])
    g, swg = engine.prep_fid(sy, wr, MAGIC, sw, npts, pw, rho, det, dts, rf=rf, step_t1=t1,
                             precision=precision)
    c, swc = O.prep_fid(sy, wr, MAGIC, sw, npts, pw, rho, det, dts, step_t1=t1, rf=rf, t0=t0)
    assert swg == swc
    err = rel(g, c)
    print(f"[prep] {name} {precision}: rel-L2 vs oracle {err:.3e}", flush=True)
    assert err < tol, (name, precision, err)
    return g


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("c64", 1e-5)])
def test_homonuclear_pulse_then_fid(precision, tol):
    """Two 1H under MAS: a 90-degree-like x pulse on Iz (the rf merges every Zeeman block of the
    preparation; the FID's blocks stay magnetisation blocks), then the FID of I+."""
    w = workloads.config1()
    sy = w.system
    dts = np.full(5, 0.5e-6)
    rf = [("1H", np.full(5, 100000.0), np.zeros(5), np.zeros(5))]
    run("config1-pulse", sy, w.wr, w.sw, 128, w.powder().subset(np.arange(0, 1154, 17)),
        sy.op("Iz"), sy.op("Ip"), dts, rf, precision, tol)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("c64", 1e-5)])
def test_cluster_pulse_with_phase_then_fid(precision, tol):
    """A 5-spin 1H cluster: a pulse with phase 30 degrees and an offset, a delay, then the FID
    (prepared rho couples every pair of the FID's magnetisation blocks)."""
    sy = workloads.small_cluster(5, 77)
    dts = np.array([0.6e-6, 0.6e-6, 0.6e-6, 2.0e-6])
    amp = np.array([90000.0, 90000.0, 90000.0, 0.0])
    rf = [("1H", amp, np.full(4, np.deg2rad(30.0)), np.array([2000.0, 2000.0, 2000.0, 0.0]))]
    run("cluster5-pulse-delay", sy, 15000.0, 15000.0 * 32, 96,
        zcw(1154).subset(np.arange(0, 1154, 61)), sy.op("Iz"), sy.op("Ip"), dts, rf, precision,
        tol)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("c64", 1e-5)])
def test_heteronuclear_13c_pulse_then_fid(precision, tol):
    """13C + two 1H: a 13C pulse (the 1H blocks unchanged), the FID detects 13C; explicit step
    start times."""
    sy = SolidSys.from_text("numspin 3\nT 13C 0\nT 1H 1\nT 1H 2\nC 200 4000 0.3 0 10 20 30\n"
                            "C 0 3000 0.6 1 40 50 60\nD -20000 0 1 0 30 40\nD -15000 0 2 0 70 10\n"
                            "D -25000 1 2 0 50 80\nJ 140 0 1\n")
    dts = np.full(4, 1.25e-6)
    t1 = 1.0e-6 + np.arange(4) * 1.25e-6
    rf = [("13C", np.full(4, 50000.0), np.zeros(4), np.zeros(4))]
    run("hetero-13C", sy, 12500.0, 12500.0 * 16, 64, zcw(1154).subset(np.arange(0, 1154, 41)),
        sy.op("Iz_0"), sy.op("Ip_0"), dts, rf, precision, tol, t1=t1)


def test_prep_rejected_on_static_plan():
    sy = workloads.config1().system
    p = engine.Plan(sy, 0.0, 0.0, 40000.0, 32)
    try:
        with pytest.raises(engine.EngineError) as e:
            p.set_prep(0.0, [1e-6], [("1H", [1e5], [0.0], [0.0])])
        assert e.value.status == 4
    finally:
        p.close()
