"""A plan run in consecutive slices of its powder (blfid_plan_run_slice) gives the whole run's
FID bit for bit (the powder sum carries its group accumulators from slice to slice in
crystallite order), also when a memory budget cuts the slices into smaller batches; slices out
of order or across a change of the plan are refused."""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import zcw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "c64"])
def test_slices_equal_the_run(precision):
    sy = workloads.small_cluster(6, 3)
    pw = zcw(1154).subset(np.arange(0, 1154, 5))
    K = len(pw)
    p = engine.Plan(sy, 12000.0, workloads.MAGIC_DEG, 12000.0 * 32, 200, precision=precision)
    try:
        p.set_operators(sy.op("Ix"), sy.op("Ip"))
        p.set_powder(pw)
        p.run()
        whole = p.fetch()
        st = p.stats()
        small = st["fixed_bytes"] + 40.5 * st["bytes_per_crystallite"]   # batches of ~40
        for cuts, budget in [([100, 101], 0.0), ([1, 64, 150], 0.0), ([37, 200], small)]:
            p.set_memory_budget(budget)
            p.run()
            ref = p.fetch()
            edges = [0] + cuts + [K]
            for a, b in zip(edges[:-1], edges[1:]):
                p.run_slice(a, b)
            got = p.fetch()
            print(f"[slices] {precision} cuts {cuts} budget {budget}: equal "
                  f"{np.array_equal(got, whole)}", flush=True)
            assert np.array_equal(ref, whole) and np.array_equal(got, whole), cuts
        with pytest.raises(engine.EngineError):
            p.run_slice(5, 10)                 # not from crystallite 0
        p.run_slice(0, 10)
        with pytest.raises(engine.EngineError):
            p.run_slice(11, 20)                # a gap
        p.run_slice(0, 10)
        p.set_operators(sy.op("Ix"), sy.op("Ip"))
        with pytest.raises(engine.EngineError):
            p.run_slice(10, K)                 # the plan changed in between
    finally:
        p.close()
