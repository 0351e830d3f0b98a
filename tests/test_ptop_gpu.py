"""GPU parity of the pulse-program path (solid-2.0's ptop(): explicit Dyson product with rf
over one rotor period, one FID point per period) against the oracle's explicit propagation.

The engine forms the period propagator U = prod_k exp(+i 2 pi tau (H_k + H_rf,k)) per
crystallite, diagonalises it and evaluates Tr(U^m rho U^-m det) in its eigenbasis; the oracle
(oracle/blfid_oracle.c, orc_ptop_fid) multiplies rho <- U rho U^H point by point.  Tolerance:
relative L2 <= 1e-9 (FP64; agreement to rounding).
"""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import Powder, zcw
from blochlib_amd.spinsys import SolidSys
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-9
MAGIC = workloads.MAGIC_DEG

HET = """numspin 3
T 13C 0
T 1H 1
T 1H 2
C 200 4000 0.3 0 10 20 30
C 0 3000 0.6 1 40 50 60
D -20000 0 1 0 30 40
D -15000 0 2 0 70 10
D -25000 1 2 0 50 80
J 140 0 1
"""


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def compare(name, sy, wr, rotor, steps, npts, pw, rho, det, rf, t0=0.0, normalize=True, tol=TOL):
    print(f"[ptop] {name}: S={sy.spins} K={len(pw)} wr={wr} steps={steps} npts={npts} "
          f"channels={len(rf)}", flush=True)
    g = engine.ptop_fid(sy, wr, rotor, steps, npts, pw, rho, det, rf=rf, t0=t0,
                        normalize=normalize)
    c = O.ptop_fid(sy, wr, rotor, steps, npts, pw, rho, det, rf=rf, t0=t0, normalize=normalize)
    err = rel_l2(g, c)
    print(f"[ptop] {name}: rel-L2 {err:.3e}", flush=True)
    assert err < tol, (name, err)
    return g, c


def xix(steps, amp, cycles, phase_shift=0.0):
    ch = workloads.phase_alternating_rf("1H", steps, amp, cycles)
    return [(n, a, p + phase_shift, o) for n, a, p, o in ch]


PW = zcw(1154).subset(np.arange(7, 1154, 163))


def test_no_rf():
    sy = SolidSys.from_text(HET)
    compare("het-no-rf", sy, 12500.0, MAGIC, 16, 40, PW, sy.op("Ix_0"), sy.op("Ip_0"), [])


@pytest.mark.parametrize("shift", [0.0, 0.7])
def test_phase_alternating_1h(shift):
    sy = SolidSys.from_text(HET)
    compare(f"het-xix-shift{shift}", sy, 12500.0, MAGIC, 32, 40, PW, sy.op("Ix_0"),
            sy.op("Ip_0"), xix(32, 60000.0, 4, shift))


def test_two_isotopes_offsets_and_zero_amplitude_steps():
    """13C channel with amplitude on half the period only (its offset counts only there) and
    two channels on 1H that the engine merges into one complex amplitude."""
    sy = SolidSys.from_text(HET)
    n = 24
    camp = np.where(np.arange(n) >= 12, 7000.0, 0.0)
    rf = xix(n, 50000.0, 3, 0.4) + [
        ("13C", camp, np.full(n, 0.3), np.full(n, 300.0)),
        ("1H", np.full(n, 8000.0), np.linspace(0, 2.0, n), np.full(n, -900.0)),
    ]
    compare("het-multi-channel", sy, 12500.0, MAGIC, n, 33, PW, sy.op("Ix_0+0.2*Iz"),
            sy.op("Ip_0"), rf, normalize=False)


@pytest.mark.parametrize("rotor,t0,steps", [(0.0, 0.0, 10), (30.0, 1.7e-5, 15)])
def test_static_rotor_start_time_odd_steps(rotor, t0, steps):
    sy = SolidSys.from_text(HET)
    compare(f"het-rotor{rotor}-n{steps}", sy, 10000.0, rotor, steps, 30, PW, sy.op("Ix"),
            sy.op("Ip"), xix(steps, 40000.0, 1) if steps % 2 == 0 else
            [("1H", np.full(steps, 40000.0), np.zeros(steps), np.zeros(steps))], t0=t0)


def test_channel_on_absent_isotope_is_ignored():
    sy = SolidSys.from_text(HET)
    rf = [("15N", np.full(8, 9000.0), np.zeros(8), np.full(8, 100.0))]
    g, _ = compare("het-absent-isotope", sy, 12500.0, MAGIC, 8, 20, PW, sy.op("Ix_0"),
                   sy.op("Ip_0"), rf)
    g0 = engine.ptop_fid(sy, 12500.0, MAGIC, 8, 20, PW, sy.op("Ix_0"), sy.op("Ip_0"))
    assert rel_l2(g, g0) < 1e-14


def test_single_spin_nutation_closed_form():
    """U = exp(+i 2 pi T a Ix): Tr(Iz_m Iz) = cos(2 pi a m / wr) / 2 and, for the y phase,
    Tr(Iz_m Ix) = -sin(2 pi a m / wr) / 2 (the + sign of the exponent)."""
    sy = SolidSys.from_text("numspin 1\nC 0 0 0 0\n")
    a, wr, n, P = 3100.0, 10000.0, 8, 40
    m = np.arange(P)
    f = engine.ptop_fid(sy, wr, MAGIC, n, P, Powder.single(), sy.op("Iz"), sy.op("Iz"),
                        rf=[("1H", a, 0.0, 0.0)], normalize=False)
    assert np.allclose(f, 0.5 * np.cos(2 * np.pi * a * m / wr), atol=1e-12)
    fy = engine.ptop_fid(sy, wr, MAGIC, n, P, Powder.single(), sy.op("Iz"), sy.op("Ix"),
                         rf=[("1H", a, np.pi / 2, 0.0)], normalize=False)
    assert np.allclose(fy, -0.5 * np.sin(2 * np.pi * a * m / wr), atol=1e-12)


def test_strong_heteronuclear_j_on_pulsed_isotope_rejected():
    sy = SolidSys.from_text(HET + "J 50 0 2 1\n")
    with pytest.raises(engine.EngineError) as e:
        engine.Plan.ptop(sy, 12500.0, MAGIC, 8, 10, rf=[("1H", 1000.0, 0.0, 0.0)])
    assert e.value.status == 4
    # without rf on either isotope of that coupling the plan is fine
    p = engine.Plan.ptop(sy, 12500.0, MAGIC, 8, 10, rf=[])
    p.close()


def test_config4_system_subset():
    """13C + 6 1H (N = 128 in two blocks of 64), 256 steps, 100 kHz phase-alternating 1H rf:
    three ZCW 1154 crystallites, 48 points."""
    w = workloads.config4(1)[0]
    rho, det = w.operators()
    pw = zcw(1154).subset(np.array([3, 500, 1000]))
    compare("config4-set0-subset", w.system, w.wr, w.rotor, w.steps, 48, pw, rho, det, w.rf)


def test_config4_full_points_invariants():
    """At the full 2048 points: fid[0] = sum(w) Tr(rho det)/Tr(det det^H) = sum(w)/2, the
    signal is linear in rho (Im_0 = Ix_0 - i Iy_0), vanishes for rho = Ip_0 (U keeps the 13C
    Zeeman blocks apart, so Tr(U^m Ip_0 U^-m Ip_0) = 0) and is bounded by Cauchy-Schwarz."""
    w = workloads.config4(2)[1]
    sy = w.system
    pw = zcw(1154).subset(np.arange(0, 1154, 72))
    det = sy.op("Ip_0")
    print(f"[ptop] config4-set1 K={len(pw)} npts={w.npts}", flush=True)
    p = engine.Plan.ptop(sy, w.wr, w.rotor, w.steps, w.npts, w.rf)
    try:
        fx = p.powder_fid(pw, sy.op("Ix_0"), det)
        fy = p.powder_fid(pw, sy.op("Iy_0"), det)
        fm = p.powder_fid(pw, sy.op("Im_0"), det)
        fp = p.powder_fid(pw, sy.op("Ip_0"), det)
    finally:
        p.close()
    sw = pw.weight.sum()
    assert abs(fx[0] - 0.5 * sw) < 1e-13
    assert rel_l2(fm, fx - 1j * fy) < 1e-12
    assert rel_l2(fm, 2.0 * fx) < 1e-12
    assert np.abs(fp).max() == 0.0
    bound = sw * np.linalg.norm(sy.op("Ix_0")) * np.linalg.norm(det) / np.linalg.norm(det) ** 2
    assert np.abs(fx).max() <= bound * (1 + 1e-12)
    assert np.all(np.isfinite(fx))


@pytest.mark.parametrize("wr,rotor", [(12500.0, MAGIC), (0.0, 0.0)])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("c64", 1e-5)])
def test_general_subsequence_unequal_steps(wr, rotor, precision, tol):
    """Pulses and delays of different lengths (blfid_plan_create_pulse), spinning or static,
    both precisions."""
    sy = SolidSys.from_text(HET)
    dts = np.array([2e-6] * 5 + [7.5e-6] * 3 + [1e-6] * 4)
    n = len(dts)
    rf = [("1H", np.r_[np.full(5, 60000.0), np.zeros(3), np.full(4, 25000.0)],
           np.r_[np.zeros(8), np.full(4, 1.1)], np.zeros(n)),
          ("13C", np.r_[np.zeros(8), np.full(4, 9000.0)], np.full(n, 0.4), np.full(n, 150.0))]
    rho, det = sy.op("Ix_0"), sy.op("Ip_0")
    print(f"[ptop] unequal steps wr={wr} {precision}", flush=True)
    g = engine.pulse_fid(sy, wr, rotor, dts, 30, PW, rho, det, rf=rf, t0=3e-6, precision=precision)
    c = O.ptop_fid(sy, wr, rotor, n, 30, PW, rho, det, rf=rf, t0=3e-6, step_dt=dts)
    err = rel_l2(g, c)
    print(f"[ptop] unequal steps: rel-L2 {err:.3e}", flush=True)
    assert err < tol
