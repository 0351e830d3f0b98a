"""ctypes front of the FP64 CPU oracle (blfid_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_powder_fid.restype = ctypes.c_int
        L.orc_powder_fid.argtypes = [ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                     ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, _d, _d, _d, _d, _d, _d, ctypes.c_int,
                                     ctypes.c_int, _d, _d]
        L.orc_ptop_fid.restype = ctypes.c_int
        L.orc_ptop_fid.argtypes = [ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                   ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _d,
                                   ctypes.c_int, _d, _d, _d, _d, _d, _d, ctypes.c_int,
                                   ctypes.c_int, _d]
        L.orc_pulse_fid.restype = ctypes.c_int
        L.orc_pulse_fid.argtypes = [ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                    ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, _d, _d, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _d, ctypes.c_int, _d, _d, _d, _d, _d, _d,
                                    ctypes.c_int, ctypes.c_int, _d]
        L.orc_prep_fid.restype = ctypes.c_int
        L.orc_prep_fid.argtypes = ([ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                    ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    _d, _d, ctypes.c_double, ctypes.c_int, _d, ctypes.c_int]
                                   + [_d] * 6 + [ctypes.c_int, ctypes.c_int, _d, _d])
        L.orc_state_apply.argtypes = ([ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                       ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _d, _d, ctypes.c_int, _d, ctypes.c_int,
                                       ctypes.c_int, _d, _d, _d, _d, ctypes.c_int])
        L.orc_state_fid.argtypes = ([ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                     ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int]
                                    + [_d] * 6 + [ctypes.c_int, ctypes.c_int, _d, _d])
        L.orc_state_pulse_fid.argtypes = ([ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                           ctypes.c_int, _d, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _d, _d, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int, _d, ctypes.c_int]
                                          + [_d] * 6 + [ctypes.c_int, ctypes.c_int, _d])
        L.orc_hamiltonian.argtypes = [ctypes.c_int, _i, ctypes.c_int, _d, ctypes.c_int, _d,
                                      ctypes.c_int, _d] + [ctypes.c_double] * 5 + [_d]
        L.orc_eig_herm_x.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_eig_normal_x.argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_expm_herm_x.argtypes = [ctypes.c_int, _d, ctypes.c_double, ctypes.c_double, _d]
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _cm(M):
    """complex128 column-major copy viewed as float64."""
    return np.asfortranarray(np.asarray(M, dtype=np.complex128)).ravel(order="F").view(np.float64)


def _sys(sy):
    iso, csa, dip, jc = sy.tables()
    iso = np.ascontiguousarray(iso, dtype=np.int32)
    return (iso, len(sy.csa), csa if csa.size else np.zeros(1), len(sy.dip),
            dip if dip.size else np.zeros(1), len(sy.jc), jc if jc.size else np.zeros(1))


def powder_fid(sy, wr, rotor_deg, sw, npts, powder, rho, det, t0=0.0, normalize=True,
               nthreads=None):
    """Powder-summed FID (complex128[npts]) and the adjusted sweep width."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    r = _cm(rho)
    d = _cm(det)
    out = np.zeros(2 * max(npts, 1), dtype=np.float64)
    swo = np.zeros(1)
    th = np.ascontiguousarray(powder.theta, np.float64)
    ph = np.ascontiguousarray(powder.phi, np.float64)
    ga = np.ascontiguousarray(powder.gamma, np.float64)
    w = np.ascontiguousarray(powder.weight, np.float64)
    nthreads = nthreads or os.cpu_count() or 1
    st = L.orc_powder_fid(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc),
                          float(wr), float(rotor_deg), float(sw), int(npts), float(t0), len(th),
                          _p(th), _p(ph), _p(ga), _p(w), _p(r), _p(d), int(bool(normalize)),
                          int(nthreads), _p(out), _p(swo))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return out[: 2 * npts].view(np.complex128).copy(), float(swo[0])


def hamiltonian(sy, theta, phi, gamma, rotor_deg, rotor_phase):
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    N = sy.N
    H = np.zeros(2 * N * N)
    L.orc_hamiltonian(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc), theta, phi,
                      gamma, rotor_deg, rotor_phase, _p(H))
    return H.view(np.complex128).reshape(N, N, order="F")


def eig_herm(A):
    N = A.shape[0]
    a = _cm(A)
    w = np.zeros(N)
    V = np.zeros(2 * N * N)
    st = lib().orc_eig_herm_x(N, _p(a), _p(w), _p(V))
    if st:
        raise RuntimeError(st)
    return w, V.view(np.complex128).reshape(N, N, order="F")


def eig_normal(A):
    N = A.shape[0]
    a = _cm(A)
    lam = np.zeros(2 * N)
    Q = np.zeros(2 * N * N)
    st = lib().orc_eig_normal_x(N, _p(a), _p(lam), _p(Q))
    if st:
        raise RuntimeError(st)
    return lam.view(np.complex128).copy(), Q.view(np.complex128).reshape(N, N, order="F")


def expm_herm(H, mul):
    N = H.shape[0]
    h = _cm(H)
    U = np.zeros(2 * N * N)
    st = lib().orc_expm_herm_x(N, _p(h), float(np.real(mul)), float(np.imag(mul)), _p(U))
    if st:
        raise RuntimeError(st)
    return U.view(np.complex128).reshape(N, N, order="F")


def rf_table(sy, rf, nsteps):
    """rf: list of (isotope name, amp[nsteps], phase_rad[nsteps], offset[nsteps]) ->
    flat [step][channel][iso id, amp, phase, offset] table."""
    ids = {}
    for n_ in sy.isotopes:
        ids.setdefault(n_, len(ids))
    tab = np.zeros((nsteps, max(len(rf), 1), 4))
    for q, (name, amp, ph, off) in enumerate(rf):
        tab[:, q, 0] = ids.get(name, -1)
        tab[:, q, 1] = np.broadcast_to(amp, nsteps)
        tab[:, q, 2] = np.broadcast_to(ph, nsteps)
        tab[:, q, 3] = np.broadcast_to(off, nsteps)
    return np.ascontiguousarray(tab.ravel()), max(len(rf), 1)


def ptop_fid(sy, wr, rotor_deg, nsteps, npts, powder, rho, det, rf=(), t0=0.0, normalize=True,
             step_dt=None, step_t1=None, nthreads=None):
    """Point-to-point FID with rf: one application of the sub-sequence per point.  The
    sub-sequence is one rotor period of nsteps equal steps, or the consecutive steps of lengths
    step_dt."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    tab, nrf = rf_table(sy, list(rf), nsteps)
    if step_dt is not None:
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(step_dt, np.float64), (nsteps,)))
    r = _cm(rho)
    d = _cm(det)
    out = np.zeros(2 * max(npts, 1))
    th = np.ascontiguousarray(powder.theta, np.float64)
    ph = np.ascontiguousarray(powder.phi, np.float64)
    ga = np.ascontiguousarray(powder.gamma, np.float64)
    w = np.ascontiguousarray(powder.weight, np.float64)
    dptr = _p(dts) if step_dt is not None else None
    if step_t1 is not None:
        t1s = np.ascontiguousarray(np.broadcast_to(np.asarray(step_t1, np.float64), (nsteps,)))
    tptr = _p(t1s) if step_t1 is not None else None
    st = L.orc_pulse_fid(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc), float(wr),
                         float(rotor_deg), int(nsteps), tptr, dptr, int(npts), float(t0), nrf,
                         _p(tab),
                         len(th), _p(th), _p(ph), _p(ga), _p(w), _p(r), _p(d),
                         int(bool(normalize)), int(nthreads or os.cpu_count() or 1), _p(out))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return out[: 2 * npts].view(np.complex128).copy()


def prep_fid(sy, wr, rotor_deg, sw, npts, powder, rho, det, step_dt, step_t1=None,
             prep_rotor_deg=None, rf=(), t0=0.0, normalize=True, nthreads=None):
    """Pulses before a gamma-COMPUTE FID: per crystallite rho <- U rho U^H with U the propagator
    of the steps (lengths step_dt, start times step_t1 or consecutive from 0, rf channels as
    ptop_fid, rotor angle prep_rotor_deg, default the FID's), then the powder FID of that rho
    from start time t0.  Returns (fid, adjusted sw)."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    dts = np.ascontiguousarray(np.asarray(step_dt, np.float64).ravel())
    nsteps = len(dts)
    tab, nrf = rf_table(sy, list(rf), nsteps)
    t1s = None if step_t1 is None else np.ascontiguousarray(np.asarray(step_t1, np.float64).ravel())
    r = _cm(rho)
    d = _cm(det)
    out = np.zeros(2 * max(npts, 1))
    swo = np.zeros(1)
    th = np.ascontiguousarray(powder.theta, np.float64)
    ph = np.ascontiguousarray(powder.phi, np.float64)
    ga = np.ascontiguousarray(powder.gamma, np.float64)
    w = np.ascontiguousarray(powder.weight, np.float64)
    st = L.orc_prep_fid(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc), float(wr),
                        float(rotor_deg), float(sw), int(npts), float(t0), nsteps,
                        _p(t1s) if t1s is not None else None, _p(dts),
                        float(rotor_deg if prep_rotor_deg is None else prep_rotor_deg), nrf,
                        _p(tab), len(th), _p(th), _p(ph), _p(ga), _p(w), _p(r), _p(d),
                        int(bool(normalize)), int(nthreads or os.cpu_count() or 1), _p(out),
                        _p(swo))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return out[: 2 * npts].view(np.complex128).copy(), float(swo[0])


# ------------------------------------------------------------------ per-crystallite rho
def _states_cm(rho_all):
    """K x N x N complex -> K consecutive column-major matrices as float64."""
    a = np.asarray(rho_all, dtype=np.complex128)
    return np.ascontiguousarray(np.transpose(a, (0, 2, 1))).view(np.float64).ravel()


def _powder_arrays(powder):
    return [np.ascontiguousarray(x, np.float64) for x in
            (powder.theta, powder.phi, powder.gamma, powder.weight)]


def state_apply(sy, wr, rotor_deg, step_dt, powder, rho_all, rf=(), step_t1=None, repeat=1,
                nthreads=None):
    """rho_k <- U_k^repeat rho_k U_k^-repeat for every crystallite k of `powder`; returns the
    new K x N x N array."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    dts = np.ascontiguousarray(np.atleast_1d(np.asarray(step_dt, np.float64)))
    n = len(dts)
    tab, nrf = rf_table(sy, list(rf), n)
    t1s = (None if step_t1 is None else
           np.ascontiguousarray(np.broadcast_to(np.asarray(step_t1, np.float64), (n,))))
    buf = _states_cm(rho_all).copy()
    th, ph, ga, _w = _powder_arrays(powder)
    st = L.orc_state_apply(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc),
                           float(wr), float(rotor_deg), n, None if t1s is None else _p(t1s),
                           _p(dts), nrf, _p(tab), int(repeat), len(th), _p(th), _p(ph), _p(ga),
                           _p(buf), int(nthreads or os.cpu_count() or 1))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    K, N = len(th), sy.N
    return np.transpose(buf.view(np.complex128).reshape(K, N, N), (0, 2, 1)).copy()


def state_fid(sy, wr, rotor_deg, sw, npts, powder, rho_all, det, t0=0.0, normalize=True,
              nthreads=None):
    """gamma-COMPUTE or static powder FID from each crystallite's own rho; (fid, sw)."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    buf = _states_cm(rho_all)
    d = _cm(det)
    out = np.zeros(2 * max(npts, 1))
    swo = np.zeros(1)
    th, ph, ga, w = _powder_arrays(powder)
    st = L.orc_state_fid(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc), float(wr),
                         float(rotor_deg), float(sw), int(npts), float(t0), len(th), _p(th),
                         _p(ph), _p(ga), _p(w), _p(buf), _p(d), int(bool(normalize)),
                         int(nthreads or os.cpu_count() or 1), _p(out), _p(swo))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return out[: 2 * npts].view(np.complex128).copy(), float(swo[0])


def state_pulse_fid(sy, wr, rotor_deg, step_dt, npts, powder, rho_all, det, rf=(),
                    step_t1=None, normalize=True, nthreads=None):
    """Point-to-point FID of a sub-sequence from each crystallite's own rho."""
    L = lib()
    iso, nc, csa, nd, dip, nj, jc = _sys(sy)
    dts = np.ascontiguousarray(np.atleast_1d(np.asarray(step_dt, np.float64)))
    n = len(dts)
    tab, nrf = rf_table(sy, list(rf), n)
    t1s = (None if step_t1 is None else
           np.ascontiguousarray(np.broadcast_to(np.asarray(step_t1, np.float64), (n,))))
    buf = _states_cm(rho_all)
    d = _cm(det)
    out = np.zeros(2 * max(npts, 1))
    th, ph, ga, w = _powder_arrays(powder)
    st = L.orc_state_pulse_fid(sy.spins, _p(iso, _i), nc, _p(csa), nd, _p(dip), nj, _p(jc),
                               float(wr), float(rotor_deg), n, None if t1s is None else _p(t1s),
                               _p(dts), int(npts), 0.0, nrf, _p(tab), len(th), _p(th), _p(ph),
                               _p(ga), _p(w), _p(buf), _p(d), int(bool(normalize)),
                               int(nthreads or os.cpu_count() or 1), _p(out))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return out[: 2 * npts].view(np.complex128).copy()
