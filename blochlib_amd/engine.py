"""ctypes binding of the engine's C ABI (include/blfid.h) and the oneFID-style front end.

The shared library is ``blochlib_amd/libblfid.so`` (built by ``python -m blochlib_amd.build``).
There is no fallback: if the library is missing or no HIP device is present, calls raise.

Front end (mirrors src/QMspins/driver/oneFID.h:52-218):

    fid = OneFID(sys, npts=256, sw=640e3, wr=10e3, rotor=54.7356)
    sig = fid.FID(powder, ro, det)             # whole powder on the GPU, oneFID.h:200-208
    sig1 = fid.FID(ro, det, theta, phi, gamma) # one crystallite, oneFID.h:210-218
    fid.sw()                                   # adjusted sweep width n*wr

Pulse programs (solid-2.0's ptop(); propogation.cc:240-304, sequencerun_funcs.cc:624-655):

    sig = ptop_fid(sys, wr, rotor, steps, npts, powder, ro, det,
                   rf=[("1H", amp[steps], phase_rad[steps], offset[steps])])
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .powder import Powder
from .spinsys import BlochLibError, SolidSys

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libblfid.so")

BLFID_OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "NOMEM", 3: "DEVICE", 4: "UNSUPPORTED", 5: "NOCONV"}
# blfid_precision: complex128 throughout, or complex64 propagators / products / correlation sums
# with the eigen-solver, phases and reductions in FP64; "bf16x3" is the complex64 form with its
# complex products as three bf16 matrix-core products each (hi*hi + hi*lo + lo*hi)
PRECISION = {"fp64": 0, "c128": 0, "c64": 1, "bf16x3": 2}

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


class EngineError(BlochLibError):
    def __init__(self, status, msg):
        super().__init__(f"blfid {STATUS.get(status, status)}: {msg}")
        self.status = status


class _System(ctypes.Structure):
    _fields_ = [("nspins", ctypes.c_int), ("isotope", _i), ("ncsa", ctypes.c_int),
                ("csa", _d), ("ndip", ctypes.c_int), ("dip", _d), ("nj", ctypes.c_int),
                ("jc", _d)]


class _RF(ctypes.Structure):
    _fields_ = [("isotope", ctypes.c_int), ("amp_hz", _d), ("phase_rad", _d),
                ("offset_hz", _d)]


_lib = None


def lib():
    """Load libblfid.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise EngineError(3, f"{LIBPATH} is not built; run python -m blochlib_amd.build")
        L = ctypes.CDLL(LIBPATH)
        vp = ctypes.c_void_p
        L.blfid_plan_create.argtypes = [ctypes.POINTER(_System), ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.POINTER(vp)]
        L.blfid_plan_create_ptop.argtypes = [ctypes.POINTER(_System), ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.POINTER(_RF),
                                             ctypes.c_int, ctypes.POINTER(vp)]
        L.blfid_plan_create_pulse.argtypes = [ctypes.POINTER(_System), ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, _d, _d,
                                              ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.POINTER(_RF), ctypes.c_int,
                                              ctypes.POINTER(vp)]
        L.blfid_plan_destroy.argtypes = [vp]
        L.blfid_plan_destroy.restype = None
        L.blfid_plan_sw.argtypes = [vp]
        L.blfid_plan_sw.restype = ctypes.c_double
        L.blfid_plan_steps.argtypes = [vp]
        L.blfid_plan_dim.argtypes = [vp]
        L.blfid_plan_set_operators.argtypes = [vp, _d, _d]
        L.blfid_plan_set_powder.argtypes = [vp, _d, _d, _d, _d, ctypes.c_int]
        L.blfid_plan_run.argtypes = [vp, ctypes.c_int, vp]
        L.blfid_plan_run_noalloc.argtypes = [vp, ctypes.c_int, vp, vp]
        L.blfid_plan_set_prep.argtypes = [vp, ctypes.c_double, ctypes.c_int, _d, _d, ctypes.c_int,
                                          ctypes.POINTER(_RF)]
        L.blfid_plan_fetch.argtypes = [vp, _d]
        L.blfid_plan_device_fid.argtypes = [vp]
        L.blfid_plan_device_fid.restype = vp
        L.blfid_powder_fid.argtypes = [vp, _d, _d, _d, _d, ctypes.c_int, _d, _d, ctypes.c_int, _d]
        L.blfid_plan_stats.argtypes = [vp, _d]
        L.blfid_plan_set_memory_budget.argtypes = [vp, ctypes.c_double]
        L.blfid_workspace_create.argtypes = [ctypes.c_double, ctypes.POINTER(vp)]
        L.blfid_workspace_destroy.argtypes = [vp]
        L.blfid_workspace_destroy.restype = None
        L.blfid_plan_set_workspace.argtypes = [vp, vp]
        L.blfid_plan_set_profiling.argtypes = [vp, ctypes.c_int]
        L.blfid_plan_phase_times.argtypes = [vp, _d, _i, _d]
        L.blfid_plan_phase_bytes.argtypes = [vp, _d]
        L.blfid_state_create.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _d,
                                         ctypes.POINTER(vp)]
        L.blfid_state_destroy.argtypes = [vp]
        L.blfid_state_destroy.restype = None
        L.blfid_state_dim.argtypes = [vp]
        L.blfid_state_size.argtypes = [vp]
        L.blfid_state_set_memory_budget.argtypes = [vp, ctypes.c_double]
        L.blfid_state_set.argtypes = [vp, _d, vp]
        L.blfid_state_load.argtypes = [vp, _d, vp]
        L.blfid_state_fetch.argtypes = [vp, _d, vp]
        L.blfid_state_apply.argtypes = [vp, ctypes.POINTER(_System), ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, _d, _d, ctypes.c_int,
                                        ctypes.POINTER(_RF), ctypes.c_int, vp]
        L.blfid_state_trace.argtypes = [vp, _d, ctypes.c_int, _d, vp]
        L.blfid_state_project.argtypes = [vp, _d, vp]
        L.blfid_plan_set_state.argtypes = [vp, vp]
        L.blfid_plan_graph_capture.argtypes = [vp, ctypes.c_int]
        L.blfid_plan_graph_launch.argtypes = [vp, vp]
        L.blfid_plan_run_slice.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.blfid_autotune.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, vp, _d]
        L.blfid_tune_get.argtypes = [_i]
        L.blfid_tune_set.argtypes = [_i]
        L.blfid_plan_tune_get.argtypes = [vp, _i]
        L.blfid_plan_tune_set.argtypes = [vp, _i]
        L.blfid_last_error.restype = ctypes.c_char_p
        L.blfid_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(st):
    if st != BLFID_OK:
        raise EngineError(st, lib().blfid_last_error().decode())


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _cm(M, N):
    M = np.asarray(M, dtype=np.complex128)
    if M.shape != (N, N):
        raise BlochLibError(f"operator must be {N}x{N}, got {M.shape}")
    return np.ascontiguousarray(M.ravel(order="F").view(np.float64))


class Workspace:
    """Device working storage shared by plans (blfid_workspace_*)."""

    def __init__(self, nbytes: float):
        h = ctypes.c_void_p()
        _check(lib().blfid_workspace_create(float(nbytes), ctypes.byref(h)))
        self.h = h
        self.nbytes = float(nbytes)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().blfid_workspace_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """One spin system + acquisition on the device (blfid_plan_*)."""

    def __init__(self, sy: SolidSys, wr: float, rotor_deg: float, sw: float, npts: int,
                 t0: float = 0.0, *, precision: str = "fp64", _ptop=None):
        L = lib()
        if precision not in PRECISION:
            raise BlochLibError(f"unknown precision {precision!r} (fp64, c64, bf16x3)")
        prec = PRECISION[precision]
        iso, csa, dip, jc = sy.tables()
        self._keep = [np.ascontiguousarray(iso, np.int32),
                      csa if csa.size else np.zeros(1), dip if dip.size else np.zeros(1),
                      jc if jc.size else np.zeros(1)]
        s = _System(sy.spins, _p(self._keep[0], _i), len(sy.csa), _p(self._keep[1]), len(sy.dip),
                    _p(self._keep[2]), len(sy.jc), _p(self._keep[3]))
        h = ctypes.c_void_p()
        if _ptop is None:
            _check(L.blfid_plan_create(ctypes.byref(s), float(wr), float(rotor_deg), float(sw),
                                       int(npts), float(t0), prec, ctypes.byref(h)))
        else:
            steps, rf, step_dt, step_t1 = _ptop
            (chans, tabs), arr = _rf_channels(sy, rf, steps)
            self._rf = (tabs, arr)
            if step_dt is None:
                _check(L.blfid_plan_create_ptop(ctypes.byref(s), float(wr), float(rotor_deg),
                                                int(steps), int(npts), float(t0), len(chans), arr,
                                                prec, ctypes.byref(h)))
            else:
                self._dt = np.ascontiguousarray(
                    np.broadcast_to(np.asarray(step_dt, np.float64), (steps,)))
                self._t1 = (None if step_t1 is None else np.ascontiguousarray(
                    np.broadcast_to(np.asarray(step_t1, np.float64), (steps,))))
                _check(L.blfid_plan_create_pulse(ctypes.byref(s), float(wr), float(rotor_deg),
                                                 int(steps),
                                                 None if self._t1 is None else _p(self._t1),
                                                 _p(self._dt), int(npts), float(t0), len(chans),
                                                 arr, prec, ctypes.byref(h)))
        self.h = h
        self.N = sy.N
        self.npts = int(npts)
        self._sys = sy

    @classmethod
    def ptop(cls, sy: SolidSys, wr: float, rotor_deg: float, steps: int, npts: int, rf=(),
             t0: float = 0.0, precision: str = "fp64") -> "Plan":
        """Rotor-synchronised pulse-program plan (blfid_plan_create_ptop).  rf: channels
        (isotope name, amp_hz, phase_rad, offset_hz), each a scalar or one value per step."""
        return cls(sy, wr, rotor_deg, 0.0, npts, t0, precision=precision,
                   _ptop=(int(steps), list(rf), None, None))

    @classmethod
    def pulse(cls, sy: SolidSys, wr: float, rotor_deg: float, step_dt, npts: int, rf=(),
              t0: float = 0.0, precision: str = "fp64", step_t1=None) -> "Plan":
        """Point-to-point plan of a general sub-sequence (blfid_plan_create_pulse): steps of
        lengths step_dt (s) starting at step_t1 (or consecutive from t0), rf channels with one
        entry per step."""
        dt = np.atleast_1d(np.asarray(step_dt, np.float64))
        return cls(sy, wr, rotor_deg, 0.0, npts, t0, precision=precision,
                   _ptop=(len(dt), list(rf), dt, step_t1))

    def set_prep(self, rotor_deg: float, step_dt, rf=(), step_t1=None):
        """Pulses applied to rho before the FID (blfid_plan_set_prep): steps of lengths step_dt
        starting at step_t1 (or consecutive from 0), rf channels (isotope, amp[steps],
        phase_rad[steps], offset[steps]); the plan's t0 should be the pulses' end.  Operators
        must be set afterwards.  An empty step list removes the preparation."""
        dt = np.ascontiguousarray(np.atleast_1d(np.asarray(step_dt, np.float64)))
        n = len(dt)
        (chans, tabs), arr = _rf_channels(self._sys, rf, n)
        t1 = (None if step_t1 is None else
              np.ascontiguousarray(np.broadcast_to(np.asarray(step_t1, np.float64), (n,))))
        self._prep = (dt, t1, tabs, arr)
        _check(lib().blfid_plan_set_prep(self.h, float(rotor_deg), n,
                                         None if t1 is None else _p(t1),
                                         _p(dt) if n else None, len(chans), arr))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().blfid_plan_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def sw(self) -> float:
        return lib().blfid_plan_sw(self.h)

    @property
    def steps(self) -> int:
        return lib().blfid_plan_steps(self.h)

    def set_operators(self, rho, det):
        self._r = _cm(rho, self.N)
        self._det = _cm(det, self.N)
        _check(lib().blfid_plan_set_operators(self.h, _p(self._r), _p(self._det)))

    def set_state(self, state: "State | None"):
        """Take each crystallite's rho from `state` (blfid_plan_set_state); None detaches.
        Operators must be set afterwards (their rho is then ignored)."""
        self._state = state
        _check(lib().blfid_plan_set_state(self.h, state.h if state is not None else None))

    def set_powder(self, pw: Powder):
        self._pw = [np.ascontiguousarray(a, np.float64) for a in
                    (pw.theta, pw.phi, pw.gamma, pw.weight)]
        _check(lib().blfid_plan_set_powder(self.h, *[_p(a) for a in self._pw], len(pw)))

    def set_workspace(self, ws: "Workspace | None"):
        self._ws = ws
        _check(lib().blfid_plan_set_workspace(self.h, ws.h if ws is not None else None))

    def set_memory_budget(self, nbytes: float):
        _check(lib().blfid_plan_set_memory_budget(self.h, float(nbytes)))

    def run(self, normalize: bool = True, stream: int | None = None):
        _check(lib().blfid_plan_run(self.h, int(bool(normalize)),
                                    ctypes.c_void_p(stream) if stream else None))

    def run_noalloc(self, fid_out_dev: int, normalize: bool = True, stream: int | None = None):
        """blfid_plan_run_noalloc: the run without device allocation, working storage in the
        library's static per-GPU scratch, FID (2 npts doubles) written to the device address
        fid_out_dev."""
        _check(lib().blfid_plan_run_noalloc(self.h, int(bool(normalize)),
                                            ctypes.c_void_p(fid_out_dev),
                                            ctypes.c_void_p(stream) if stream else None))

    def run_slice(self, c_begin: int, c_end: int, normalize: bool = True,
                  stream: int | None = None):
        """Crystallites [c_begin, c_end) (blfid_plan_run_slice): slices from 0 to K in order give
        run()'s FID."""
        _check(lib().blfid_plan_run_slice(self.h, int(c_begin), int(c_end), int(bool(normalize)),
                                          ctypes.c_void_p(stream) if stream else None))

    def graph_capture(self, normalize: bool = True):
        """Record the whole run as a HIP graph (blfid_plan_graph_capture); nothing runs."""
        _check(lib().blfid_plan_graph_capture(self.h, int(bool(normalize))))

    def graph_launch(self, stream: int | None = None):
        """Run the recorded graph (blfid_plan_graph_launch); then fetch() as after run()."""
        _check(lib().blfid_plan_graph_launch(self.h, ctypes.c_void_p(stream) if stream else None))

    def fetch(self) -> np.ndarray:
        out = np.zeros(2 * self.npts)
        _check(lib().blfid_plan_fetch(self.h, _p(out)))
        return out.view(np.complex128).copy()

    def device_fid_ptr(self) -> int:
        return lib().blfid_plan_device_fid(self.h) or 0

    def stats(self) -> dict:
        out = np.zeros(6)
        _check(lib().blfid_plan_stats(self.h, _p(out)))
        return {"flops_per_cs": out[0], "dense_flops_per_cs": out[1], "blocks": int(out[2]),
                "max_block": int(out[3]), "bytes_per_crystallite": out[4],
                "fixed_bytes": out[5]}

    PHASES = ("exp", "dyson", "eig", "bmul", "qr", "corr", "reduce", "static", "expdyson")
    KERNELS = ("k_exp", "k_dyson", "k_eig", "k_bmul", "k_qr", "k_corr", "k_reduce",
               "k_static_fid", "k_expdyson")
    NSLOT = 72  # BLFID_NSLOT

    def set_profiling(self, on: bool = True):
        _check(lib().blfid_plan_set_profiling(self.h, int(bool(on))))

    def phase_times(self) -> dict:
        """{kernel name: {ms, launches, flops, bytes}} for every slot that ran, e.g.
        'k_exp<5>' (bytes: algorithmic HBM bytes)."""
        ms = np.zeros(self.NSLOT)
        n = np.zeros(self.NSLOT, dtype=np.int32)
        fl = np.zeros(self.NSLOT)
        _check(lib().blfid_plan_phase_times(self.h, _p(ms), _p(n, _i), _p(fl)))
        by = np.zeros(self.NSLOT)
        _check(lib().blfid_plan_phase_bytes(self.h, _p(by)))
        out = {}
        for slot in range(self.NSLOT):
            if n[slot]:
                ph, M = divmod(slot, 8)
                if ph == 2:   # eigen-solver size classes
                    name = "k_eig<" + ("16", "32", "48", "64", "80", "blk")[M - 1] + ">"
                else:
                    name = self.KERNELS[ph] + (f"<{M}>" if M else "")
                out[name] = {"ms": float(ms[slot]), "launches": int(n[slot]),
                             "flops": float(fl[slot]), "bytes": float(by[slot])}
        return out

    def tuning(self) -> dict:
        """The plan's launch choices (blfid_plan_tune_get) as a dict (see tuning_to_dict)."""
        v = np.zeros(TUNE_LEN, np.int32)
        _check(lib().blfid_plan_tune_get(self.h, _p(v, _i)))
        return tuning_to_dict(v)

    def set_tuning(self, tuning):
        """Set the plan's launch choices (blfid_plan_tune_set) from a dict or TUNE_LEN ints."""
        v = tuning_to_array(tuning)
        _check(lib().blfid_plan_tune_set(self.h, _p(v, _i)))

    def powder_fid(self, pw: Powder, rho, det, normalize=True) -> np.ndarray:
        self.set_operators(rho, det)
        self.set_powder(pw)
        self.run(normalize)
        return self.fetch()


TUNE_LEN = 16          # BLFID_TUNE_LEN
TUNE_FAMILIES = ("exp", "dyson", "eig", "bmul", "qr", "corr", "reduce", "static", "expdyson")


def tuning_to_dict(v) -> dict:
    """BLFID_TUNE_LEN ints -> {"method": {M: "fused"|"two"}, "corr": 0..2, "grid": {family: cap}
    (cap 0: the default), "qr_flip": {M: the other Q-product form}}."""
    v = [int(x) for x in v]
    return {"method": {M: ("two" if v[M - 1] else "fused") for M in range(1, 6)},
            "corr": v[5], "grid": {f: v[6 + i] for i, f in enumerate(TUNE_FAMILIES)},
            "qr_flip": {M: bool((v[15] >> (M - 1)) & 1) for M in range(1, 5)}}


def tuning_to_array(t) -> np.ndarray:
    if isinstance(t, dict):
        v = np.zeros(TUNE_LEN, np.int32)
        for M, m in t.get("method", {}).items():
            v[int(M) - 1] = 1 if m in ("two", 1) else 0
        v[5] = int(t.get("corr", 0))
        for i, f in enumerate(TUNE_FAMILIES):
            v[6 + i] = int(t.get("grid", {}).get(f, 0))
        v[15] = sum(1 << (int(M) - 1) for M, on in t.get("qr_flip", {}).items() if on)
        return v
    v = np.ascontiguousarray(t, np.int32)
    if v.shape != (TUNE_LEN,):
        raise ValueError(f"a tuning is {TUNE_LEN} ints")
    return v


def tuning() -> dict:
    """The current device's process-wide launch choices (blfid_tune_get)."""
    v = np.zeros(TUNE_LEN, np.int32)
    _check(lib().blfid_tune_get(_p(v, _i)))
    return tuning_to_dict(v)


def set_tuning(t):
    """Set the current device's process-wide launch choices (blfid_tune_set): plans created
    afterwards take them (e.g. a tuning saved as JSON by an earlier process)."""
    v = tuning_to_array(t)
    _check(lib().blfid_tune_set(_p(v, _i)))


def autotune(plans, crystallites: int = 64, stream: int | None = None) -> dict:
    """Time the launch candidates on the first `crystallites` crystallites of each plan and set
    the fastest on every plan and on the process table (blfid_autotune); returns the chosen
    tuning with the summed kernel time of each candidate pass ("pass_ms").  The plans' last FIDs
    are overwritten: run them again before fetching."""
    plans = list(plans)
    arr = (ctypes.c_void_p * len(plans))(*[p.h.value for p in plans])
    rep = np.zeros(7)
    _check(lib().blfid_autotune(arr, len(plans), int(crystallites),
                                ctypes.c_void_p(stream) if stream else None, _p(rep)))
    out = plans[0].tuning()
    out["pass_ms"] = [float(x) for x in rep]
    return out


def _system(sy: SolidSys):
    """(_System, arrays it points into) for one SolidSys."""
    iso, csa, dip, jc = sy.tables()
    keep = [np.ascontiguousarray(iso, np.int32), csa if csa.size else np.zeros(1),
            dip if dip.size else np.zeros(1), jc if jc.size else np.zeros(1)]
    s = _System(sy.spins, _p(keep[0], _i), len(sy.csa), _p(keep[1]), len(sy.dip), _p(keep[2]),
                len(sy.jc), _p(keep[3]))
    return s, keep


class State:
    """Per-crystallite density operators of one powder in device memory (blfid_state_*): the
    rho that solid-2.0 carries through a pulses{} program for every powder orientation
    (sequencerun.h:268, subSequence::propogate sequencerun.cc:164-196)."""

    def __init__(self, nspins: int, pw: Powder):
        self._pw = [np.ascontiguousarray(a, np.float64) for a in
                    (pw.theta, pw.phi, pw.gamma, pw.weight)]
        h = ctypes.c_void_p()
        _check(lib().blfid_state_create(int(nspins), len(pw), *[_p(a) for a in self._pw],
                                        ctypes.byref(h)))
        self.h = h
        self.N = 1 << int(nspins)
        self.K = len(pw)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().blfid_state_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, rho):
        """Every crystallite's rho = rho (ro(op))."""
        r = _cm(rho, self.N)
        _check(lib().blfid_state_set(self.h, _p(r), None))

    def load(self, rho_all):
        """Crystallite k's rho = rho_all[k] (K x N x N)."""
        a = np.asarray(rho_all, dtype=np.complex128)
        if a.shape != (self.K, self.N, self.N):
            raise BlochLibError(f"states must be {self.K}x{self.N}x{self.N}, got {a.shape}")
        buf = np.ascontiguousarray(np.transpose(a, (0, 2, 1))).view(np.float64)
        _check(lib().blfid_state_load(self.h, _p(buf), None))

    def fetch(self) -> np.ndarray:
        """K x N x N array of every crystallite's rho."""
        buf = np.zeros(2 * self.K * self.N * self.N)
        _check(lib().blfid_state_fetch(self.h, _p(buf), None))
        return np.transpose(buf.view(np.complex128).reshape(self.K, self.N, self.N),
                            (0, 2, 1)).copy()

    def set_memory_budget(self, nbytes: float):
        _check(lib().blfid_state_set_memory_budget(self.h, float(nbytes)))

    def apply(self, sy: SolidSys, wr: float, rotor_deg: float, step_dt, rf=(), step_t1=None,
              repeat: int = 1):
        """rho <- U^repeat rho U^-repeat, U the propagator of the steps (lengths step_dt,
        starts step_t1 or consecutive from 0, rf channels (isotope, amp[n], phase_rad[n],
        offset[n])) at every crystallite's orientation."""
        dt = np.ascontiguousarray(np.atleast_1d(np.asarray(step_dt, np.float64)))
        n = len(dt)
        t1 = (None if step_t1 is None else
              np.ascontiguousarray(np.broadcast_to(np.asarray(step_t1, np.float64), (n,))))
        (chans, tabs), arr = _rf_channels(sy, rf, n)
        s, keep = _system(sy)
        _check(lib().blfid_state_apply(self.h, ctypes.byref(s), float(wr), float(rotor_deg), n,
                                       None if t1 is None else _p(t1), _p(dt), len(chans), arr,
                                       int(repeat), None))

    def project(self, op):
        """rho_k <- Tr(rho_k op^H) op / Tr(op op^H) for every crystallite (cycler(op))."""
        o = _cm(op, self.N)
        _check(lib().blfid_state_project(self.h, _p(o), None))

    def trace(self, det, normalize: bool = True) -> complex:
        """sum_k w_k Tr(rho_k det) [/ Tr(det det^H)]."""
        d = _cm(det, self.N)
        out = np.zeros(2)
        _check(lib().blfid_state_trace(self.h, _p(d), int(bool(normalize)), _p(out), None))
        return complex(out[0], out[1])


def state_fid(sy, wr, rotor_deg, sw, npts, pw, state: State, det, t0=0.0, normalize=True,
              precision="fp64"):
    """FID(powder, ro_k, det) from each crystallite's current rho (a gamma-COMPUTE or static
    plan reading `state`); returns (fid, adjusted sw)."""
    p = Plan(sy, wr, rotor_deg, sw, npts, t0, precision=precision)
    try:
        p.set_state(state)
        f = p.powder_fid(pw, det, det, normalize)
        return f, p.sw
    finally:
        p.close()


def state_pulse_fid(sy, wr, rotor_deg, step_dt, npts, pw, state: State, det, rf=(),
                    step_t1=None, normalize=True):
    """Point-to-point FID of a sub-sequence from each crystallite's current rho: point m is
    sum_k w_k Tr(U^m rho_k U^-m det) [/ Tr(det det^H)]."""
    p = Plan.pulse(sy, wr, rotor_deg, step_dt, npts, rf, 0.0, "fp64", step_t1)
    try:
        p.set_state(state)
        return p.powder_fid(pw, det, det, normalize)
    finally:
        p.close()


def powder_fid(sy, wr, rotor_deg, sw, npts, pw, rho, det, t0=0.0, normalize=True,
               precision="fp64"):
    """FID(powder&, ro, det) on the GPU; returns (fid, adjusted sw)."""
    p = Plan(sy, wr, rotor_deg, sw, npts, t0, precision=precision)
    try:
        f = p.powder_fid(pw, rho, det, normalize)
        return f, p.sw
    finally:
        p.close()


def _rf_channels(sy, rf, steps):
    """((channels, tables), blfid_rf array) of rf channels (isotope name, amp, phase_rad,
    offset; scalars or one value per step); the array points into the tables."""
    ids = {}
    for n_ in sy.isotopes:
        ids.setdefault(n_, len(ids))
    tabs, chans = [], []
    for name, amp, ph, off in rf:
        t = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (steps,)))
             for v in (amp, ph, off)]
        tabs.append(t)
        chans.append(_RF(ids.get(name, -1), _p(t[0]), _p(t[1]), _p(t[2])))
    arr = (_RF * max(1, len(chans)))(*chans)
    return (chans, tabs), arr


def prep_fid(sy, wr, rotor_deg, sw, npts, pw, rho, det, step_dt, rf=(), step_t1=None,
             t0=None, prep_rotor_deg=None, normalize=True, precision="fp64"):
    """Pulses before a gamma-COMPUTE FID (solid-2.0: pulses, then a plain fid()): per
    crystallite rho <- U rho U^H with U the steps' propagator, then the powder FID from start
    time t0 (default: the end of the steps).  Returns (fid, adjusted sw)."""
    dt = np.atleast_1d(np.asarray(step_dt, np.float64))
    if t0 is None:
        t0 = float(dt.sum()) if step_t1 is None else float(np.asarray(step_t1)[-1] + dt[-1])
    p = Plan(sy, wr, rotor_deg, sw, npts, t0, precision=precision)
    try:
        p.set_prep(rotor_deg if prep_rotor_deg is None else prep_rotor_deg, dt, rf, step_t1)
        return p.powder_fid(pw, rho, det, normalize), p.sw
    finally:
        p.close()


def pulse_fid(sy, wr, rotor_deg, step_dt, npts, pw, rho, det, rf=(), t0=0.0, normalize=True,
              precision="fp64", step_t1=None):
    """Powder-summed point-to-point FID of a general sub-sequence (steps of lengths step_dt,
    starting at step_t1 or consecutive from t0), one FID point per application of the whole
    sub-sequence."""
    p = Plan.pulse(sy, wr, rotor_deg, step_dt, npts, rf, t0, precision, step_t1)
    try:
        return p.powder_fid(pw, rho, det, normalize)
    finally:
        p.close()


def ptop_fid(sy, wr, rotor_deg, steps, npts, pw, rho, det, rf=(), t0=0.0, normalize=True,
             precision="fp64"):
    """Powder-summed ptop FID of a one-rotor-period pulse program on the GPU (one point per
    period): sum_k w_k Tr(rho_m det) [/ Tr(det det^H)], rho_{m+1} = U rho_m U^H."""
    p = Plan.ptop(sy, wr, rotor_deg, steps, npts, rf, t0, precision)
    try:
        return p.powder_fid(pw, rho, det, normalize)
    finally:
        p.close()


class OneFID:
    """oneFID<SolidSys, hmatrix> front end (src/QMspins/driver/oneFID.h:52-218)."""

    def __init__(self, sys: SolidSys, npts: int, sw: float = 20000.0, wr: float = 0.0,
                 rotb: float = 0.0, startT: float = 0.0):
        self.function_ = sys
        self._npts = int(npts)
        self.sw_ = float(sw)
        self.wr_ = float(wr)
        self.rotb_ = float(rotb)
        self.startT_ = float(startT)

    # setters/getters of oneFID.h:130-177
    def setFunction(self, sys): self.function_ = sys
    def setWr(self, v): self.wr_ = float(v)
    def setRotorSpeed(self, v): self.wr_ = float(v)
    def setRotorAngle(self, v): self.rotb_ = float(v)
    def setRotorBeta(self, v): self.rotb_ = float(v)
    def setSw(self, v): self.sw_ = float(v)
    def setSweepWidth(self, v): self.sw_ = float(v)
    def setSize(self, n): self._npts = int(n)
    def setNpts(self, n): self._npts = int(n)
    def setStartTime(self, t): self.startT_ = float(t)
    def wr(self): return self.wr_
    def rotorAngle(self): return self.rotb_
    def sw(self): return self.sw_
    def npts(self): return self._npts
    def size(self): return self._npts
    def startTime(self): return self.startT_

    def FID(self, *args, normalize=False):
        """FID(powder, ro, det) or FID(ro, det, theta=0, phi=0, gamma=0).  Like the library
        call (oneFID.h:396-402) the powder sum is NOT divided by Tr(det det^H) unless
        normalize=True (solid-2.0's doFID does divide, sequencerun_funcs.cc:698-702)."""
        if self._npts <= 0:
            raise BlochLibError("Fid size was '0' please use 'npts(int)' set set it")
        if isinstance(args[0], Powder):
            pw, ro, det = args[0], args[1], args[2]
        else:
            ro, det = args[0], args[1]
            th, ph, ga = (list(args[2:5]) + [0.0, 0.0, 0.0])[:3]
            pw = Powder.single(th, ph, ga, 1.0)
        p = Plan(self.function_, self.wr_, self.rotb_, self.sw_, self._npts, self.startT_)
        try:
            f = p.powder_fid(pw, ro, det, normalize)
            self.sw_ = p.sw
            return f
        finally:
            p.close()
