"""Launch autotuner (blfid_autotune, blfid_tune_*): every grid cap and correlation variant gives
the default run's FID bit for bit; the two-kernel propagator form of the complex64 path stays
within the complex64 tolerance of the fused form and of the FP64 form; a tuning pass on several
plans sets one table on each plan and on the process table, which plans created later inherit;
bad tunings are refused."""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import zcw

pytestmark = pytest.mark.gpu
MAGIC = workloads.MAGIC_DEG
ZERO = np.zeros(engine.TUNE_LEN, np.int32)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make(S, seed, precision, K=96, npts=384):
    sy = workloads.small_cluster(S, seed)
    p = engine.Plan(sy, 12000.0, MAGIC, 12000.0 * 64, npts, precision=precision)
    p.set_operators(sy.op("Ix"), sy.op("Ip"))
    p.set_powder(zcw(1154).subset(np.arange(K) * 11 % 1154))
    return p


@pytest.fixture(autouse=True)
def reset_process_table():
    engine.set_tuning(ZERO)
    yield
    engine.set_tuning(ZERO)


def test_forced_choices_keep_the_fid():
    p = make(7, 3, "c64")
    d = make(7, 3, "fp64")
    try:
        p.run()
        base = p.fetch()
        d.run()
        ref = d.fetch()
        for name, t in [("corr 1", {"corr": 1}), ("corr 2", {"corr": 2}),
                        ("grid 1024", {"grid": {f: 1024 for f in engine.TUNE_FAMILIES}}),
                        ("grid 8192", {"grid": {f: 8192 for f in engine.TUNE_FAMILIES}})]:
            p.set_tuning(t)
            p.run()
            got = p.fetch()
            print(f"[tune] c64 {name}: equal {np.array_equal(got, base)}", flush=True)
            assert np.array_equal(got, base), name
            d.set_tuning(t)
            d.run()
            assert np.array_equal(d.fetch(), ref), name
        p.set_tuning({"method": {M: "two" for M in range(1, 6)}})
        p.run()
        two = p.fetch()
        print(f"[tune] c64 two-kernel form: vs fused {rel(two, base):.2e}, vs FP64 "
              f"{rel(two, ref):.2e}, fused vs FP64 {rel(base, ref):.2e}", flush=True)
        assert rel(two, ref) < 1e-5 and rel(base, ref) < 1e-5
        assert p.tuning()["method"][3] == "two"
        p.set_tuning({"qr_flip": {M: True for M in range(1, 5)}})
        p.run()
        flip = p.fetch()
        print(f"[tune] c64 other Q-product forms: vs default {rel(flip, base):.2e}, vs FP64 "
              f"{rel(flip, ref):.2e}", flush=True)
        assert rel(flip, ref) < 1e-5 and rel(flip, base) < 2e-6
        assert p.tuning()["qr_flip"] == {1: True, 2: True, 3: True, 4: True}
    finally:
        p.close()
        d.close()


@pytest.mark.parametrize("precision", ["c64", "fp64"])
def test_autotune_sets_every_plan_and_the_process_table(precision):
    plans = [make(5, 11, precision), make(7, 12, precision, K=64)]
    try:
        before = []
        for p in plans:
            p.run()
            before.append(p.fetch())
        t = engine.autotune(plans, crystallites=32)
        print(f"[tune] {precision}: {t}", flush=True)
        assert len(t["pass_ms"]) == 7 and all(x > 0 for x in t["pass_ms"])
        table = engine.tuning()
        for p in plans:
            assert p.tuning() == table
        with pytest.raises(engine.EngineError):
            plans[0].fetch()                   # the tuning runs overwrote the FID
        for p, b in zip(plans, before):
            p.run()
            got = p.fetch()
            if precision == "fp64" or all(m == "fused" for m in table["method"].values()):
                assert np.array_equal(got, b)
            else:
                assert rel(got, b) < 1e-5
        q = make(6, 13, precision, K=16)
        try:
            assert q.tuning() == table         # a new plan takes the process table
        finally:
            q.close()
    finally:
        for p in plans:
            p.close()


def test_bad_tunings_are_refused():
    p = make(4, 5, "c64", K=8)
    try:
        for bad in [dict(corr=3), {"grid": {"qr": -1}}]:
            with pytest.raises((engine.EngineError, ValueError)):
                p.set_tuning(bad)
        v = ZERO.copy()
        v[0] = 2
        with pytest.raises(engine.EngineError):
            engine.set_tuning(v)
        v = ZERO.copy()
        v[15] = 16
        with pytest.raises(engine.EngineError):
            p.set_tuning(v)
        assert p.tuning() == engine.tuning_to_dict(ZERO)
    finally:
        p.close()
