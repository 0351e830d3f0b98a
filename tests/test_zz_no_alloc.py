"""blfid_plan_run_noalloc (the run that allocates no device memory, working storage in the
library's static per-GPU scratch) against blfid_plan_run through the C ABI: the same FID bit for
bit, on two calls in a row (the second with a smaller input), and a typed NOMEM, before any
launch, for a plan whose one crystallite exceeds the scratch."""
import ctypes

import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import zcw

pytestmark = pytest.mark.gpu


def _hip():
    L = ctypes.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return L


def _run_both(p, pw, rho, det, hip):
    p.set_operators(rho, det)
    p.set_powder(pw)
    ref = p.powder_fid(pw, rho, det)
    npts = len(ref)
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 16 * npts) == 0
    try:
        p.run_noalloc(dev.value, True)
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(2 * npts)
        assert hip.hipMemcpy(out.ctypes.data, dev.value, 16 * npts, 2) == 0  # device to host
    finally:
        hip.hipFree(dev)
    return out.view(np.complex128), ref


@pytest.mark.parametrize("precision", ["c64", "fp64"])
def test_noalloc_equals_run_two_calls(precision):
    hip = _hip()
    w = next(x for x in workloads.config2() if x.meta["spins"] == 6)
    rho, det = w.operators()
    p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts, precision=precision)
    try:
        for K in (400, 37):  # second call: a smaller input on the same plan
            pw = w.powder().subset(np.arange(K) * 7 % w.powder_size)
            print(f"[noalloc] {w.name} {precision} K={K}", flush=True)
            got, ref = _run_both(p, pw, rho, det, hip)
            assert np.array_equal(got, ref), (K, np.abs(got - ref).max())
    finally:
        p.close()
    # a different, smaller plan right after: the scratch keeps the last call's contents
    c1 = workloads.config1()
    rho1, det1 = c1.operators()
    p1 = engine.Plan(c1.system, c1.wr, c1.rotor, c1.sw, c1.npts, precision=precision)
    try:
        print(f"[noalloc] config1 {precision} K=100", flush=True)
        got, ref = _run_both(p1, c1.powder().subset(np.arange(100) * 11), rho1, det1, hip)
        assert np.array_equal(got, ref)
    finally:
        p1.close()


def test_noalloc_rejects_plans_beyond_the_scratch():
    w = workloads.config3()  # 10 spins: one crystallite needs ~0.3 GB of working storage
    rho, det = w.operators()
    p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts, precision="c64")
    try:
        p.set_operators(rho, det)
        p.set_powder(w.powder().subset([0, 1]))
        print("[noalloc] config3 beyond the scratch", flush=True)
        with pytest.raises(engine.EngineError) as e:
            p.run_noalloc(1 << 20)  # never dereferenced: the size check comes first
        assert e.value.status == 2
        assert "no-alloc scratch" in str(e.value)
    finally:
        p.close()
