"""Pins the FP64 CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * linear algebra   -> numpy/scipy (LAPACK) on random and degenerate inputs
  * Hamiltonian      -> an independent numpy restatement (Kronecker operators, elementary
                        rotations) and geometric closed forms (axial P2, magic-angle average)
  * static path      -> the reference's stored output examples/solid-2.0/examples/static/simpSTA
  * gamma-COMPUTE    -> explicit step propagation (oracle/bruteforce.py), rosym and general rho,
                        even and odd n (integer-division scale), wrap-around of the rotor period
  * closed forms     -> fid[0] and a single-spin MAS FID from a scalar phase sum
  * states           -> per-crystallite propagation against scipy expm steps, composition and
                        repeat 0; FIDs from per-crystallite rho against the uniform-rho and
                        prepared-rho FIDs, and the pulse FID shifted by one point
  * ptop with rf     -> explicit numpy propagation with the rf Hamiltonian built from Kronecker
                        operators, on- and off-resonance nutation closed forms (propagator
                        sign, phase sense, offsets, the zero-amplitude channel rule)
"""
import numpy as np
import pytest
import scipy.linalg

from blochlib_amd.powder import Powder
from blochlib_amd.spinsys import SolidSys
from oracle import bruteforce as B

MAGIC = float(np.degrees(np.arccos(1 / np.sqrt(3))))


def rand_herm(rng, N):
    A = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    return (A + A.conj().T) / 2


@pytest.mark.parametrize("N", [1, 2, 3, 5, 8, 16, 33])
def test_eig_herm_matches_lapack(oracle, N):
    rng = np.random.default_rng(N)
    A = rand_herm(rng, N)
    w, V = oracle.eig_herm(A)
    assert np.allclose(np.sort(w), np.linalg.eigvalsh(A), atol=1e-12 * N)
    assert np.allclose(V.conj().T @ V, np.eye(N), atol=1e-13 * N)
    assert np.allclose(V @ np.diag(w) @ V.conj().T, A, atol=1e-12 * N)


def test_eig_herm_degenerate(oracle):
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12)))
    d = np.array([1, 1, 1, -2, -2, 0.5, 3, 3, 3, 3, 0, 0], float)
    A = Q @ np.diag(d) @ Q.conj().T
    w, V = oracle.eig_herm(A)
    assert np.allclose(np.sort(w), np.sort(d), atol=1e-12)
    assert np.allclose(V @ np.diag(w) @ V.conj().T, A, atol=1e-12)


@pytest.mark.parametrize("N", [1, 2, 4, 7, 16, 40])
def test_eig_normal_unitary(oracle, N):
    rng = np.random.default_rng(100 + N)
    U = scipy.linalg.expm(-1j * 3.0 * rand_herm(rng, N))
    lam, W = oracle.eig_normal(U)
    assert np.allclose(W.conj().T @ W, np.eye(N), atol=1e-13 * N)
    assert np.allclose(U @ W, W * lam, atol=1e-12 * N)
    ref = np.linalg.eigvals(U)
    assert np.allclose(np.sort_complex(lam), np.sort_complex(ref), atol=1e-12 * N)


def test_eig_normal_degenerate_blocks(oracle):
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.normal(size=(10, 10)) + 1j * rng.normal(size=(10, 10)))
    ph = np.exp(1j * np.array([0.3, 0.3, 0.3, -1.0, -1.0, 2.0, np.pi, np.pi, 0.0, 0.0]))
    U = Q @ np.diag(ph) @ Q.conj().T
    lam, W = oracle.eig_normal(U)
    assert np.allclose(W.conj().T @ W, np.eye(10), atol=1e-12)
    assert np.allclose(U @ W, W * lam, atol=1e-12)


@pytest.mark.parametrize("N", [2, 4, 9, 32])
def test_expm_matches_scipy(oracle, N):
    rng = np.random.default_rng(200 + N)
    H = rand_herm(rng, N)
    mul = -1j * 0.37
    U = oracle.expm_herm(H, mul)
    assert np.allclose(U, scipy.linalg.expm(mul * H), atol=1e-12)


SYS3 = """numspin 3
T 1H 0
T 1H 1
T 13C 2
C 1000 5000 0.5 0 10 20 30
C -1000 3000 0.2 1 40 50 60
C 300 7000 0.9 2
C 0 4000 0 2 5 6 7
C 250 0 0 1
D -20000 0 1 0 30 40
D 5000 1 2 10 70 20
J 140 0 2
J 50 0 1 1
"""


def test_hamiltonian_matches_independent_numpy(oracle):
    sy = SolidSys.from_text(SYS3)
    rng = np.random.default_rng(11)
    for _ in range(6):
        th, ph, ga = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
        for rot, al in [(MAGIC, rng.uniform(0, 6.3)), (0.0, 0.0), (23.0, rng.uniform(0, 6.3))]:
            H1 = oracle.hamiltonian(sy, th, ph, ga, rot, al)
            H2 = B.hamiltonian(sy, th, ph, ga, rot, al)
            assert np.allclose(H1, H2, atol=1e-9), (rot, al)
            assert np.allclose(H1, H1.conj().T)


def test_axial_csa_is_delta_p2(oracle):
    """eta = 0: zz = delta * P2(cos angle(tensor axis, B0)), computed from unit vectors."""
    sy = SolidSys.from_text("numspin 1\nC 0 7000 0 0 20 35 50\n")
    rng = np.random.default_rng(5)
    for _ in range(5):
        th, ph, ga = rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
        H = oracle.hamiltonian(sy, th, ph, ga, 0.0, 0.0)
        # tensor axis in the crystal frame = third column of E (E maps PAS -> crystal)
        E = B.euler(*np.radians([20, 35, 50]))
        axis_c = E[:, 2]
        # lab z in the crystal frame = third row of the powder rotation
        z_c = B.euler(ph, th, ga)[2]
        c = axis_c @ z_c
        zz = 7000 * (3 * c * c - 1) / 2
        assert np.isclose(H[0, 0].real, 0.5 * zz, atol=1e-9)


def test_magic_angle_rotor_average_vanishes(oracle):
    """Average of the anisotropic term over a rotor cycle is P2(cos beta_r) * (axis part):
    it vanishes at the magic angle for every orientation and eta."""
    sy = SolidSys.from_text("numspin 1\nC 0 9000 0.7 0 12 77 140\n")
    phases = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    for th, ph, ga in [(0.3, 1.1, 2.0), (2.0, 4.0, 0.5)]:
        avg = np.mean([oracle.hamiltonian(sy, th, ph, ga, MAGIC, a)[0, 0].real for a in phases])
        assert abs(avg) < 1e-9
        avg30 = np.mean([oracle.hamiltonian(sy, th, ph, ga, 30.0, a)[0, 0].real for a in phases])
        assert abs(avg30) > 1.0


def test_static_fid_matches_reference_output(oracle):
    """examples/solid-2.0/examples/static/static-basic.sim -> simpSTA (6 significant digits,
    produced by the reference's complex64 diagonal path)."""
    sy = SolidSys.from_text("numspin 2\nT 1H 0\nT 1H 1\nC 5000  2134 0 0\nC -5000 2789 0.5 1\n")
    fid, sw = oracle.powder_fid(sy, 0.0, 0.0, 40000.0, 512, Powder.read("rep2000"),
                                sy.op("Ix"), sy.op("Ip"), normalize=True)
    ref = np.loadtxt("tests/golden/simpSTA.txt")
    fr = ref[:, 4] + 1j * ref[:, 5]
    assert sw == 40000.0
    assert np.abs(fid - fr).max() < 2e-6
    assert np.linalg.norm(fid - fr) / np.linalg.norm(fr) < 1e-5


@pytest.mark.parametrize("sw", [80000.0, 70000.0, 120000.0])
@pytest.mark.parametrize("rho_name", ["Ix", "Iz+0.3*Ix_0"])
def test_gamma_compute_equals_explicit_stepping(oracle, sw, rho_name):
    sy = SolidSys.from_text(SYS3)
    rho, det = sy.op(rho_name), sy.op("Ip")
    wr, npts = 10000.0, 27
    for th, ph, ga in [(0.7, 1.9, 2.3), (2.9, 0.2, 5.5)]:
        f1, swo = oracle.powder_fid(sy, wr, MAGIC, sw, npts, Powder.single(th, ph, ga), rho, det,
                                    normalize=False, nthreads=1)
        f2 = B.gamma_compute_equivalent(sy, th, ph, ga, MAGIC, wr, sw, npts, rho, det)
        assert swo == int(np.floor(sw / wr + 0.5)) * wr
        assert np.abs(f1 - f2).max() < 1e-11 * max(1.0, np.abs(f2).max())


def test_period_phase_is_reference_not_forward(oracle):
    """The reference applies whole rotor periods through U_T^-1 (see bruteforce docstring):
    the forward-propagated signal agrees only before the first wrap."""
    sy = SolidSys.from_text(SYS3)
    rho, det = sy.op("Ix"), sy.op("Ip")
    f1, _ = oracle.powder_fid(sy, 10000.0, MAGIC, 80000.0, 20, Powder.single(0.7, 1.9, 2.3), rho,
                              det, normalize=False, nthreads=1)
    fw = B.gamma_compute_equivalent(sy, 0.7, 1.9, 2.3, MAGIC, 10000.0, 80000.0, 20, rho, det,
                                    forward=True)
    assert np.isclose(f1[0], fw[0])
    assert np.abs(f1 - fw).max() > 1e-3


def test_fid0_closed_form_and_powder_sum(oracle):
    """fid[0] = sum(w) * (Tr(D^2) + Tr(D D^H)) * 2 / Tr(D D^H) for rho = (D+D^H)/2 (= 2 for
    D = Fp), and the powder sum is linear in the weights."""
    sy = SolidSys.from_text(SYS3)
    pw = Powder.read("ZCW_3_1154").subset(np.arange(0, 1154, 97))
    rho, det = sy.op("Ix"), sy.op("Ip")
    f, _ = oracle.powder_fid(sy, 10000.0, MAGIC, 40000.0, 9, pw, rho, det, normalize=True)
    assert np.isclose(f[0], 2.0 * pw.weight.sum(), atol=1e-12)
    pw2 = Powder(pw.theta, pw.phi, pw.gamma, 2.5 * pw.weight)
    f2, _ = oracle.powder_fid(sy, 10000.0, MAGIC, 40000.0, 9, pw2, rho, det, normalize=True,
                              nthreads=3)
    assert np.allclose(f2, 2.5 * f, atol=1e-12)


def test_single_spin_mas_closed_form(oracle):
    """One spin, CSA only: H(t) = w(t) Iz commutes, so each start index i gives
    2 * 0.5 * exp(i 2 pi tau * phase sum) with the period part entering as U_T^-1."""
    sy = SolidSys.from_text("numspin 1\nC 800 6000 0.4 0 10 40 70\n")
    wr, sw, npts = 5000.0, 40000.0, 21
    n = 8
    tau = 1 / wr / n
    th, ph, ga = 1.1, 0.4, 2.2
    w = np.array([2 * oracle.hamiltonian(sy, th, ph, ga, MAGIC, 2 * np.pi * wr * (k + 0.5) * tau)[0, 0].real
                  for k in range(n)])
    cum = np.concatenate([[0.0], np.cumsum(w)])  # A_k phase for |+><-|
    fid = np.zeros(npts, complex)
    for i in range(n):
        for t in range(npts):
            M = (i + t) // n
            j = (i + t) % n
            phase = cum[j + 1] - M * cum[n] - cum[i + 1]
            fid[t] += np.exp(2j * np.pi * tau * phase)
    fid *= 2 * 0.5 / (n // 2)
    f1, _ = oracle.powder_fid(sy, wr, MAGIC, sw, npts, Powder.single(th, ph, ga), sy.op("Ix"),
                              sy.op("Ip"), normalize=False, nthreads=1)
    assert np.abs(f1 - fid).max() < 1e-12


# ------------------------------------------------------------------ ptop with rf pulses
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


def _xix(nsteps, amp, period_segments=4):
    seg = nsteps // period_segments
    ph = np.array([0.0 if (k // seg) % 2 == 0 else np.pi for k in range(nsteps)])
    return [("1H", np.full(nsteps, amp), ph, np.zeros(nsteps))]


@pytest.mark.parametrize("phase_shift", [0.0, 0.7])
def test_ptop_matches_explicit_numpy(oracle, phase_shift):
    sy = SolidSys.from_text(HET)
    rf = _xix(16, 50000.0)
    # 13C channel: amplitude only on the second half of the period (its offset counts only
    # where the amplitude is non-zero, pulsedata.cc:114)
    camp = np.where(np.arange(16) >= 8, 7000.0, 0.0)
    rf = [(n, a, p + phase_shift, o) for n, a, p, o in rf] + [("13C", camp, np.full(16, 0.3),
                                                               np.full(16, 300.0))]
    rho, det = sy.op("Ix_0"), sy.op("Ip_0")
    for th, ph, ga in [(0.4, 1.2, 2.0), (2.2, 5.0, 0.3)]:
        f1 = oracle.ptop_fid(sy, 12500.0, MAGIC, 16, 12, Powder.single(th, ph, ga), rho, det, rf=rf,
                             normalize=False)
        f2 = B.ptop_equivalent(sy, th, ph, ga, MAGIC, 12500.0, 16, 12, rho, det, rf=rf)
        assert np.abs(f1 - f2).max() < 1e-11


def test_ptop_single_spin_nutation_closed_form(oracle):
    """One spin, no interactions, on-resonance rf amp a along x for the whole period:
    Tr(Iz(t) Iz) = 1/2 cos(2 pi a m T) with U = exp(+i 2 pi tau a Ix)."""
    sy = SolidSys.from_text("numspin 1\nC 0 0 0 0\n")
    a, wr, n = 3100.0, 10000.0, 8
    f = oracle.ptop_fid(sy, wr, MAGIC, n, 9, Powder.single(), sy.op("Iz"), sy.op("Iz"),
                        rf=[("1H", np.full(n, a), np.zeros(n), np.zeros(n))], normalize=False)
    m = np.arange(9)
    assert np.allclose(f, 0.5 * np.cos(2 * np.pi * a * m / wr), atol=1e-13)
    # phase 90 deg (Iy): Iz -> nutation about y has the same Iz trace; Ix detect flips sign
    fy = oracle.ptop_fid(sy, wr, MAGIC, n, 9, Powder.single(), sy.op("Iz"), sy.op("Ix"),
                         rf=[("1H", np.full(n, a), np.full(n, np.pi / 2), np.zeros(n))],
                         normalize=False)
    # exp(+i phi Iy) Iz exp(-i phi Iy) = Iz cos phi - Ix sin phi  -> Tr(. Ix) = -sin(phi)/2
    # (the + sign of the propagator exponent turns the nutation the other way round)
    assert np.allclose(fy, -0.5 * np.sin(2 * np.pi * a * m / wr), atol=1e-13)


def test_ptop_offset_closed_forms(oracle):
    """One spin, rf amp a and offset o for the whole period: U = exp(+i 2 pi T (a Ix + o Iz))
    precesses about n = (a, 0, o)/w_eff, so Tr(Iz_m Iz) = 1/2 (n_z^2 + (1 - n_z^2) cos(2 pi
    w_eff m T)).  A channel with a == 0 contributes nothing, its offset included."""
    sy = SolidSys.from_text("numspin 1\nC 0 0 0 0\n")
    a, o, wr, n = 2300.0, 1700.0, 10000.0, 8
    m = np.arange(11)
    f = oracle.ptop_fid(sy, wr, MAGIC, n, 11, Powder.single(), sy.op("Iz"), sy.op("Iz"),
                        rf=[("1H", np.full(n, a), np.zeros(n), np.full(n, o))], normalize=False)
    weff = np.hypot(a, o)
    nz2 = (o / weff) ** 2
    assert np.allclose(f, 0.5 * (nz2 + (1 - nz2) * np.cos(2 * np.pi * weff * m / wr)), atol=1e-13)
    f0 = oracle.ptop_fid(sy, wr, MAGIC, n, 11, Powder.single(), sy.op("Ix"), sy.op("Ix"),
                         rf=[("1H", np.zeros(n), np.zeros(n), np.full(n, o))], normalize=False)
    assert np.allclose(f0, 0.5, atol=1e-14)


@pytest.mark.parametrize("wr", [12500.0, 0.0])
def test_unequal_steps_match_explicit_numpy(oracle, wr):
    """A general sub-sequence: consecutive steps of different lengths (pulses and delays split
    into substeps), also without spinning."""
    sy = SolidSys.from_text(HET)
    dts = np.array([2e-6] * 5 + [7.5e-6] * 3 + [1e-6] * 4)
    n = len(dts)
    amp = np.r_[np.full(5, 60000.0), np.zeros(3), np.full(4, 25000.0)]
    ph = np.r_[np.zeros(5), np.zeros(3), np.full(4, 1.1)]
    rf = [("1H", amp, ph, np.zeros(n)), ("13C", np.r_[np.zeros(8), np.full(4, 9000.0)],
                                        np.full(n, 0.4), np.full(n, 150.0))]
    rho, det = sy.op("Ix_0"), sy.op("Ip_0")
    for th, ph_, ga in [(0.4, 1.2, 2.0), (2.2, 5.0, 0.3)]:
        f1 = oracle.ptop_fid(sy, wr, MAGIC if wr else 0.0, n, 10, Powder.single(th, ph_, ga), rho,
                             det, rf=rf, t0=3e-6, normalize=False, step_dt=dts)
        f2 = B.ptop_equivalent(sy, th, ph_, ga, MAGIC if wr else 0.0, wr, n, 10, rho, det, rf=rf,
                               t0=3e-6, step_dt=dts)
        assert np.abs(f1 - f2).max() < 1e-11


def test_explicit_step_start_times_match_explicit_numpy(oracle):
    """Two sub-sequences each timed from 0 (non-consecutive step start times)."""
    sy = SolidSys.from_text(HET)
    dts = np.array([3e-6, 3e-6, 5e-6, 2e-6, 2e-6, 2e-6])
    t1s = np.array([0.0, 3e-6, 6e-6, 0.0, 2e-6, 4e-6])
    n = len(dts)
    rf = [("1H", np.r_[np.full(3, 40000.0), np.full(3, 70000.0)], np.r_[np.zeros(3), np.full(3, 0.5)],
           np.zeros(n))]
    rho, det = sy.op("Ix_0"), sy.op("Ip_0")
    f1 = oracle.ptop_fid(sy, 9000.0, MAGIC, n, 12, Powder.single(0.9, 2.1, 0.4), rho, det, rf=rf,
                         normalize=False, step_dt=dts, step_t1=t1s)
    f2 = B.ptop_equivalent(sy, 0.9, 2.1, 0.4, MAGIC, 9000.0, n, 12, rho, det, rf=rf, step_dt=dts,
                           step_t1=t1s)
    f3 = oracle.ptop_fid(sy, 9000.0, MAGIC, n, 12, Powder.single(0.9, 2.1, 0.4), rho, det, rf=rf,
                         normalize=False, step_dt=dts)
    assert np.abs(f1 - f2).max() < 1e-11
    assert np.abs(f1 - f3).max() > 1e-6   # the start times matter under spinning


@pytest.mark.parametrize("case", ["homo-x-pulse", "hetero-13C-pulse-and-delay"])
def test_prep_fid_equals_explicit_propagation(oracle, case):
    """Pulses before a gamma-COMPUTE fid() (sequencerun_funcs.cc:693-703): the oracle's
    prepared-rho FID equals, per crystallite, U_prep rho U_prep^H from scipy expm steps (+i sign,
    rf on the pulsed isotope) followed by the explicit-stepping gamma-COMPUTE equivalent from the
    pulses' end time."""
    if case == "homo-x-pulse":
        sy = SolidSys.from_text("numspin 2\nT 1H 0\nT 1H 1\nC 1000 5000 0.5 0 10 20 30\n"
                                "C -1000 4000 0.3 1 40 50 60\nD -15000 0 1 0 30 40\n")
        dts = np.full(3, 0.8e-6)
        rf = [("1H", np.full(3, 100000.0), np.zeros(3), np.zeros(3))]
        rho, det = sy.op("Iz"), sy.op("Ip")
    else:
        sy = SolidSys.from_text("numspin 2\nT 13C 0\nT 1H 1\nC 200 4000 0.3 0 10 20 30\n"
                                "C 0 3000 0.6 1 40 50 60\nD -20000 0 1 0 30 40\nJ 140 0 1\n")
        dts = np.array([1.25e-6, 1.25e-6, 2.0e-6])
        rf = [("13C", np.array([80000.0, 80000.0, 0.0]), np.array([0.0, 0.3, 0.0]),
               np.array([500.0, 500.0, 0.0]))]
        rho, det = sy.op("Iz_0"), sy.op("Ip_0")
    wr, sw, npts = 10000.0, 80000.0, 20
    t0 = float(dts.sum())
    for th, ph, ga in [(0.7, 1.9, 2.3), (1.2, 0.4, 0.0)]:
        f1, _ = oracle.prep_fid(sy, wr, MAGIC, sw, npts, Powder.single(th, ph, ga), rho, det,
                                step_dt=dts, rf=rf, t0=t0, normalize=False)
        f2 = B.prep_equivalent(sy, th, ph, ga, MAGIC, wr, sw, npts, rho, det, dts, rf=rf, t0_fid=t0)
        assert np.abs(f1 - f2).max() < 1e-11 * max(1.0, np.abs(f2).max()), (case, th)


# ------------------------------------------------------------------ per-crystallite rho (states)
def _state_case():
    sy = SolidSys.from_text("numspin 3\nT 1H 0\nT 1H 1\nT 13C 2\nC 1000 5000 0.5 0 10 20 30\n"
                            "C -800 4000 0.3 1 40 50 60\nC 300 6000 0.2 2 5 15 25\n"
                            "D -15000 0 1 0 30 40\nD 3000 0 2 10 20 30\nJ 60 1 2 0\n")
    pw = Powder(np.array([0.7, 1.2, 2.1]), np.array([1.9, 0.4, 3.0]), np.array([2.3, 0.0, 1.1]),
                np.array([0.5, 0.3, 0.2]))
    dts = np.array([0.8e-6, 0.8e-6, 1.5e-6, 2.0e-6])
    t1s = np.array([3e-6, 3.8e-6, 4.6e-6, 6.1e-6])
    rf = [("1H", np.array([90000.0, 90000.0, 0.0, 50000.0]), np.array([0.0, 0.5, 0.0, 1.0]),
           np.array([300.0, 300.0, 0.0, -200.0])),
          ("13C", np.array([0.0, 60000.0, 60000.0, 0.0]), np.array([0.0, 0.2, 0.2, 0.0]),
           np.zeros(4))]
    return sy, pw, dts, t1s, rf


@pytest.mark.parametrize("wr,repeat", [(10000.0, 1), (10000.0, 3), (0.0, 2)])
def test_state_apply_equals_explicit_propagators(oracle, wr, repeat):
    """rho_k <- U_k^m rho_k U_k^-m with U_k from scipy expm steps (+i sign, rf per isotope,
    explicit start times) at each crystallite's orientation; every crystallite starts from its
    own random Hermitian rho so that a mix-up of crystallites or a dropped U^H fails."""
    sy, pw, dts, t1s, rf = _state_case()
    rng = np.random.default_rng(5)
    rho = np.stack([rand_herm(rng, sy.N) for _ in range(len(pw))])
    got = oracle.state_apply(sy, wr, MAGIC, dts, pw, rho, rf=rf, step_t1=t1s, repeat=repeat)
    for k in range(len(pw)):
        U = B.pulse_propagator(sy, pw.theta[k], pw.phi[k], pw.gamma[k], MAGIC, wr, len(dts), rf,
                               0.0, dts, t1s)
        Um = np.linalg.matrix_power(U, repeat)
        want = Um @ rho[k] @ Um.conj().T
        assert np.abs(got[k] - want).max() < 1e-12 * np.abs(want).max(), k
        assert np.isclose(np.trace(got[k]), np.trace(rho[k]))


def test_state_apply_repeat_zero_and_composition(oracle):
    sy, pw, dts, t1s, rf = _state_case()
    rng = np.random.default_rng(9)
    rho = np.stack([rand_herm(rng, sy.N) for _ in range(len(pw))])
    same = oracle.state_apply(sy, 8000.0, MAGIC, dts, pw, rho, rf=rf, step_t1=t1s, repeat=0)
    assert np.array_equal(same, rho)
    once = oracle.state_apply(sy, 8000.0, MAGIC, dts, pw, rho, rf=rf, step_t1=t1s)
    twice = oracle.state_apply(sy, 8000.0, MAGIC, dts, pw, once, rf=rf, step_t1=t1s)
    both = oracle.state_apply(sy, 8000.0, MAGIC, dts, pw, rho, rf=rf, step_t1=t1s, repeat=2)
    assert np.abs(twice - both).max() < 1e-13 * np.abs(both).max()


@pytest.mark.parametrize("wr", [10000.0, 0.0])
def test_state_fid_of_a_uniform_rho_is_the_powder_fid(oracle, wr):
    """Every crystallite at the same rho: the per-crystallite FID is the plain powder FID
    (gamma-COMPUTE when spinning, the static FID when not), rosym rho included."""
    sy, pw, _, _, _ = _state_case()
    det = sy.op("Ip_0+Ip_1")
    for rho in (sy.op("Ix_0+Ix_2"), 0.5 * (det + det.conj().T)):
        rho_all = np.stack([rho] * len(pw))
        f1, sw1 = oracle.state_fid(sy, wr, MAGIC, 80000.0, 32, pw, rho_all, det, t0=2e-6)
        f2, sw2 = oracle.powder_fid(sy, wr, MAGIC, 80000.0, 32, pw, rho, det, t0=2e-6)
        assert sw1 == sw2 and np.abs(f1 - f2).max() < 1e-13 * np.abs(f2).max()


def test_state_fid_after_apply_is_the_prepared_fid(oracle):
    """rho_k from state_apply of the pulse steps, then the FID from the pulses' end: the
    prepared-rho FID (itself pinned against explicit propagation above)."""
    sy, pw, dts, t1s, rf = _state_case()
    rho, det = sy.op("Iz_0+Iz_1"), sy.op("Ip_2")
    rho_all = oracle.state_apply(sy, 10000.0, MAGIC, dts, pw, np.stack([rho] * len(pw)), rf=rf,
                                 step_t1=t1s)
    t0 = float(t1s[-1] + dts[-1])
    f1, _ = oracle.state_fid(sy, 10000.0, MAGIC, 80000.0, 24, pw, rho_all, det, t0=t0)
    f2, _ = oracle.prep_fid(sy, 10000.0, MAGIC, 80000.0, 24, pw, rho, det, step_dt=dts,
                            step_t1=t1s, rf=rf, t0=t0)
    assert np.abs(f1 - f2).max() < 1e-12 * np.abs(f2).max()


def test_state_pulse_fid_continues_the_pulse_fid(oracle):
    """The pulse-program FID from rho_k = U_k rho U_k^H is the plain one shifted by a point;
    from a uniform rho it is the plain one."""
    sy, pw, dts, t1s, rf = _state_case()
    rho, det = sy.op("Ix"), sy.op("Ip_0")
    base = oracle.ptop_fid(sy, 12000.0, MAGIC, len(dts), 17, pw, rho, det, rf=rf, step_dt=dts,
                           step_t1=t1s)
    uni = oracle.state_pulse_fid(sy, 12000.0, MAGIC, dts, 17, pw, np.stack([rho] * len(pw)),
                                 det, rf=rf, step_t1=t1s)
    assert np.abs(uni - base).max() < 1e-13 * np.abs(base).max()
    moved = oracle.state_apply(sy, 12000.0, MAGIC, dts, pw, np.stack([rho] * len(pw)), rf=rf,
                               step_t1=t1s)
    f = oracle.state_pulse_fid(sy, 12000.0, MAGIC, dts, 16, pw, moved, det, rf=rf, step_t1=t1s)
    assert np.abs(f - base[1:]).max() < 1e-12 * np.abs(base).max()
