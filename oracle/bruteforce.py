"""Second, independent oracle for tiny systems (numpy/scipy, pure restatement from the physics).

TEST INFRASTRUCTURE ONLY.

* The Hamiltonian is rebuilt from Kronecker-product spin operators and from the rotations
  written as products of elementary axis rotations, not from the C oracle's code:
      cartPowder = Rz(gamma) Ry(theta) Rz(phi)        (== rotationMatrix3D, space_ten.cc:578-618)
      cartSpin   = Rx(beta_r) Rz(alpha)                (== rotationMatrix3D(a,b), :620-652)
  and the interaction definitions of csa.cc:289-329, dip.cc:171-207, spin_ten.cc:431-465.
* The gamma-COMPUTE FID (compute.h:308-461) is replaced by explicit products of step
  propagators exp(-i 2 pi tau H(t_mid)) (scipy.linalg.expm): for every start index i of the
  rotor period (the n gamma shifts 2*pi*i/n) and every time t, Tr(D V rho_eff V^H) is summed
  with the propagator V(i,t) spelled out in ``gamma_compute_equivalent``.  The reference's
  result is (1/(n div ff)) * that sum, where rho_eff = rho and ff = 1 in general, and
  rho_eff = D + D^H, ff = 2 when rho == (D + D^H)/2 exactly (compute.h:75-83, 381-382,
  458-459).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

from blochlib_amd.spinsys import single_spin_op

DEG = np.pi / 180.0


def Rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


def Ry(b):
    c, s = np.cos(b), np.sin(b)
    return np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])


def Rx(b):
    c, s = np.cos(b), np.sin(b)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def euler(a, b, g):
    return Rz(g) @ Ry(b) @ Rz(a)


def lab_rotation(theta, phi, gamma, rotor_rad, alpha):
    P = euler(phi, theta, gamma)
    if rotor_rad == 0.0:
        return P
    return Rx(rotor_rad) @ Rz(alpha) @ P


def hamiltonian(sy, theta, phi, gamma, rotor_deg, alpha):
    S, N = sy.spins, sy.N
    R = lab_rotation(theta, phi, gamma, rotor_deg * DEG, alpha)
    H = np.zeros((N, N), dtype=np.complex128)
    op = lambda c, i: single_spin_op(S, c, i)
    beta0 = rotor_deg == 0.0
    for iso, dl, eta, on, a, b, g in sy.csa:
        Iz = op("z", on)
        if iso != 0:
            H += -iso * Iz
            if dl == 0.0:
                continue
        if beta0 and eta == 0.0 and dl == 0.0:
            continue
        E = euler(a * DEG, b * DEG, g * DEG)
        C = E @ np.diag([dl * (eta - 1) / 2, -dl * (eta + 1) / 2, dl]) @ E.T
        H += (R @ C @ R.T)[2, 2] * Iz
    for dl, i, j, a, b, g in sy.dip:
        if beta0 and dl == 0.0:
            continue
        E = euler(a * DEG, b * DEG, g * DEG)
        s6 = np.sqrt(6.0)
        C = E @ np.diag([-dl * s6 / 2, -dl * s6 / 2, dl * s6]) @ E.T
        zz = (R @ C @ R.T)[2, 2]
        if sy.isotopes[i] == sy.isotopes[j]:
            T = (2 * op("z", i) @ op("z", j) - op("x", i) @ op("x", j) - op("y", i) @ op("y", j))
            H += zz / s6 * T
        else:
            H += zz * np.sqrt(2.0 / 3.0) * op("z", i) @ op("z", j)
    for J, i, j, strong in sy.jc:
        T = op("z", i) @ op("z", j)
        if strong:
            T = T + op("x", i) @ op("x", j) + op("y", i) @ op("y", j)
        H += J * T
    return H


def step_propagators(sy, theta, phi, gamma, rotor_deg, wr, n, t0=0.0):
    tau = 1.0 / wr / n
    Us = []
    for k in range(n):
        tmid = t0 + (k + 0.5) * tau
        H = hamiltonian(sy, theta, phi, gamma, rotor_deg, 2 * np.pi * wr * tmid)
        Us.append(scipy.linalg.expm(-2j * np.pi * tau * H))
    return Us


def gamma_compute_equivalent(sy, theta, phi, gamma, rotor_deg, wr, sw, npts, rho, det, t0=0.0,
                             forward=False):
    """Reference gamma-COMPUTE FID of one crystallite, by explicit stepping.

    With A_k = U_k ... U_1 (A_0 = 1) and U_T = A_n, and Qs[k], rotmp[k] built from A_{k+1}
    (compute.h:378-383), the reference's phase bookkeeping (compute.h:406-457:
    exp(-i ind2 tau w_rs) with ind2 = p-n on wrap, times z^t) equals, for start index
    i = 0..n-1 and time t, the propagator

        V(i, t) = A_{((i+t) mod n)+1}  U_T^(-M)  A_{i+1}^H,      M = floor((i+t)/n),

    i.e. the in-period steps run forward but whole rotor periods enter through U_T^(-1).  A
    physically forward propagation would use U_T^(+M) (``forward=True``); both agree for
    i + t < n.  The parity target is the reference, so ``forward=False`` is the default.
    """
    n = int(np.floor(sw / wr + 0.5))
    Us = step_propagators(sy, theta, phi, gamma, rotor_deg, wr, n, t0)
    rosym = np.array_equal(rho, 0.5 * (det + det.conj().T))
    ff = 2 if rosym else 1
    rho_eff = det + det.conj().T if rosym else rho
    N = sy.N
    A = [np.eye(N, dtype=np.complex128)]
    for k in range(n):
        A.append(Us[k] @ A[-1])
    UT = A[n]
    UTinv = UT.conj().T
    fid = np.zeros(npts, dtype=np.complex128)
    for i in range(n):
        for t in range(npts):
            M = (i + t) // n
            P = np.linalg.matrix_power(UT if forward else UTinv, M)
            V = A[(i + t) % n + 1] @ P @ A[i + 1].conj().T
            fid[t] += np.trace(det @ V @ rho_eff @ V.conj().T)
    return fid / (n // ff)


def static_fid(sy, theta, phi, gamma, rotor_deg, sw, npts, rho, det):
    """Static FID by explicit propagation with exp(-i 2 pi dt H) (oneFID.h:292-336, without the
    reference's |A|^2 <= 1e-10 pruning)."""
    H = hamiltonian(sy, theta, phi, gamma, rotor_deg, 0.0)
    U = scipy.linalg.expm(-2j * np.pi * H / sw)
    fid = np.zeros(npts, dtype=np.complex128)
    r = rho.astype(np.complex128)
    for t in range(npts):
        fid[t] = np.trace(det @ r)
        r = U @ r @ U.conj().T
    return fid


def rf_hamiltonian(sy, rf_k):
    """PulseData::Pulse (pulsedata.cc:110-125) for one step: rf_k = [(isotope, amp, phase,
    offset)]: amp (cos(ph) Ix_i + sin(ph) Iy_i) + offset Iz_i on spins of that isotope; a
    channel with amp == 0 adds nothing, not even its offset."""
    S = sy.spins
    H = np.zeros((sy.N, sy.N), dtype=np.complex128)
    for name, amp, ph, off in rf_k:
        if amp == 0.0:
            continue
        for i in range(S):
            if sy.isotopes[i] == name:
                H += amp * (np.cos(ph) * single_spin_op(S, "x", i) + np.sin(ph) * single_spin_op(S, "y", i))
                H += off * single_spin_op(S, "z", i)
    return H


def pulse_propagator(sy, theta, phi, gamma, rotor_deg, wr, nsteps, rf=(), t0=0.0, step_dt=None,
                     step_t1=None):
    """U = prod_k exp(+i 2 pi tau_k (H_k + H_rf,k)) over one rotor period of equal steps or the
    given steps (propogation.cc:240-304; U *= X is U = X U), from scipy.linalg.expm."""
    dts = (np.full(nsteps, 1.0 / wr / nsteps) if step_dt is None
           else np.broadcast_to(np.asarray(step_dt, float), (nsteps,)))
    starts = (t0 + np.concatenate([[0.0], np.cumsum(dts)[:-1]]) if step_t1 is None
              else np.broadcast_to(np.asarray(step_t1, float), (nsteps,)))
    U = np.eye(sy.N, dtype=np.complex128)
    for k in range(nsteps):
        tau = dts[k]
        tmid = starts[k] + 0.5 * tau
        H = hamiltonian(sy, theta, phi, gamma, rotor_deg, 2 * np.pi * wr * tmid)
        H = H + rf_hamiltonian(sy, [(n_, np.broadcast_to(a, nsteps)[k], np.broadcast_to(p, nsteps)[k],
                                     np.broadcast_to(o, nsteps)[k]) for n_, a, p, o in rf])
        U = scipy.linalg.expm(2j * np.pi * tau * H) @ U
    return U


def prep_equivalent(sy, theta, phi, gamma, rotor_deg, wr, sw, npts, rho, det, step_dt, rf=(),
                    step_t1=None, t0_fid=0.0, prep_rotor_deg=None):
    """Pulses before a gamma-COMPUTE FID (sequencerun_funcs.cc:693-703) for one crystallite:
    rho <- U rho U^H with U the pulse steps' propagator, then the explicit-stepping equivalent of
    the reference's gamma-COMPUTE FID of that rho from start time t0_fid."""
    dts = np.asarray(step_dt, float)
    U = pulse_propagator(sy, theta, phi, gamma,
                         rotor_deg if prep_rotor_deg is None else prep_rotor_deg, wr, len(dts), rf,
                         0.0, dts, step_t1)
    rc = U @ rho.astype(np.complex128) @ U.conj().T
    return gamma_compute_equivalent(sy, theta, phi, gamma, rotor_deg, wr, sw, npts, rc, det,
                                    t0=t0_fid)


def ptop_equivalent(sy, theta, phi, gamma, rotor_deg, wr, nsteps, npts, rho, det, rf=(), t0=0.0,
                    step_dt=None, step_t1=None):
    """One crystallite: U = prod_k exp(+i 2 pi tau_k (H_k + H_rf,k)) (propogation.cc:240-304,
    note the + sign), fid[m] = Tr(rho_m det), rho_{m+1} = U rho_m U^H (sequencerun.cc:164-196).
    Steps are one rotor period of equal length, or consecutive steps of lengths step_dt."""
    U = pulse_propagator(sy, theta, phi, gamma, rotor_deg, wr, nsteps, rf, t0, step_dt, step_t1)
    r = rho.astype(np.complex128)
    fid = np.zeros(npts, dtype=np.complex128)
    for m in range(npts):
        fid[m] = np.trace(r @ det)
        r = U @ r @ U.conj().T
    return fid
