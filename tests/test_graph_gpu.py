"""A plan's run recorded as a HIP graph (blfid_plan_graph_capture / blfid_plan_graph_launch):
every launch of the recorded graph equals the eager run bit for bit, with other eager work and
synchronisations between the launches; a graph of a plan that reads a state sees the state's
current content; any change to the plan invalidates the graph until it is captured again."""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import zcw

pytestmark = pytest.mark.gpu
MAGIC = workloads.MAGIC_DEG


@pytest.mark.parametrize("precision", ["fp64", "c64"])
def test_graph_launches_equal_the_eager_run(precision):
    w = workloads.config1()
    sy = w.system
    pw = w.powder().subset(np.arange(0, 1154, 7))
    other = workloads.small_cluster(5, 77)
    print(f"[graph] config1 {precision} K={len(pw)}", flush=True)
    p = engine.Plan(sy, w.wr, w.rotor, w.sw, 256, precision=precision)
    q = engine.Plan(other, 15000.0, MAGIC, 15000.0 * 32, 96, precision=precision)
    try:
        p.set_operators(sy.op("Ix"), sy.op("Ip"))
        p.set_powder(pw)
        p.run()
        eager = p.fetch()
        p.graph_capture()
        q.set_operators(other.op("Ix"), other.op("Ip"))
        q.set_powder(zcw(1154).subset(np.arange(0, 1154, 61)))
        for rep in range(3):
            q.run()                      # other eager work between the launches
            _ = q.fetch()
            p.graph_launch()
            got = p.fetch()
            print(f"[graph] launch {rep}: equal {np.array_equal(got, eager)}", flush=True)
            assert np.array_equal(got, eager), rep
            p.run()                      # an eager run of the same plan in between
            assert np.array_equal(p.fetch(), eager)
        # a change to the plan invalidates the graph
        p.set_powder(pw.subset(np.arange(10)))
        with pytest.raises(engine.EngineError) as e:
            p.graph_launch()
        assert e.value.status == 1
        p.graph_capture()
        p.graph_launch()
        small = p.fetch()
        p.run()
        assert np.array_equal(small, p.fetch())
    finally:
        p.close()
        q.close()


def test_graph_reads_the_current_state():
    sy = workloads.small_cluster(4, 5)
    pw = zcw(1154).subset(np.arange(0, 1154, 37))
    dts = np.full(3, 0.5e-6)
    rf = [("1H", np.full(3, 100000.0), np.zeros(3), np.zeros(3))]
    st = engine.State(sy.spins, pw)
    p = engine.Plan(sy, 12000.0, MAGIC, 12000.0 * 16, 64)
    try:
        st.set(sy.op("Iz"))
        p.set_state(st)
        p.set_operators(sy.op("Ip"), sy.op("Ip"))
        p.set_powder(pw)
        p.graph_capture()
        for rep in range(3):
            st.apply(sy, 12000.0, MAGIC, dts, rf=rf)
            p.run()
            eager = p.fetch()
            p.graph_launch()
            got = p.fetch()
            print(f"[graph] state launch {rep}: equal {np.array_equal(got, eager)}", flush=True)
            assert np.array_equal(got, eager), rep
    finally:
        p.close()
        st.close()


def test_graph_capture_refusals():
    w = workloads.config1()
    p = engine.Plan(w.system, w.wr, w.rotor, w.sw, 32)
    ws = engine.Workspace(64e6)
    try:
        with pytest.raises(engine.EngineError):
            p.graph_capture()                       # no operators yet
        p.set_operators(w.system.op("Ix"), w.system.op("Ip"))
        p.set_powder(w.powder().subset(np.arange(5)))
        with pytest.raises(engine.EngineError):
            p.graph_launch()                        # nothing captured
        p.set_profiling(True)
        with pytest.raises(engine.EngineError) as e:
            p.graph_capture()
        assert e.value.status == 4
        p.set_profiling(False)
        p.set_workspace(ws)
        with pytest.raises(engine.EngineError) as e:
            p.graph_capture()
        assert e.value.status == 4
    finally:
        p.close()
        ws.close()
