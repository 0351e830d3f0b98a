"""Per-crystallite density operators on the device (blfid_state_*), the rho that solid-2.0
carries through a pulses{} program for every powder orientation (SequenceRun's ro,
sequencerun.h:268; subSequence::propogate, sequencerun.cc:164-196), and the FIDs that start from
it (blfid_plan_set_state).  GPU through the C ABI against the oracle's state functions (pinned in
test_oracle.py): propagation, traces and FIDs in FP64 to 1e-10 relative, complex64 FIDs to 1e-5;
the batch split of a propagation is checked bitwise."""
import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import Powder, zcw
from blochlib_amd.spinsys import SolidSys
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MAGIC = workloads.MAGIC_DEG


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def rand_herm_stack(rng, K, N):
    A = rng.normal(size=(K, N, N)) + 1j * rng.normal(size=(K, N, N))
    return (A + np.conj(np.transpose(A, (0, 2, 1)))) / 2


def case3():
    """13C + two 1H with every interaction kind, rf on both isotopes, explicit start times."""
    sy = SolidSys.from_text("numspin 3\nT 1H 0\nT 1H 1\nT 13C 2\nC 1000 5000 0.5 0 10 20 30\n"
                            "C -800 4000 0.3 1 40 50 60\nC 300 6000 0.2 2 5 15 25\n"
                            "D -15000 0 1 0 30 40\nD 3000 0 2 10 20 30\nJ 60 1 2 0\n")
    dts = np.array([0.8e-6, 0.8e-6, 1.5e-6, 2.0e-6])
    t1s = np.array([3e-6, 3.8e-6, 4.6e-6, 6.1e-6])
    rf = [("1H", np.array([90000.0, 90000.0, 0.0, 50000.0]), np.array([0.0, 0.5, 0.0, 1.0]),
           np.array([300.0, 300.0, 0.0, -200.0])),
          ("13C", np.array([0.0, 60000.0, 60000.0, 0.0]), np.array([0.0, 0.2, 0.2, 0.0]),
           np.zeros(4))]
    return sy, zcw(1154).subset(np.arange(0, 1154, 23)), dts, t1s, rf


def test_set_load_fetch_round_trip():
    sy, pw, *_ = case3()
    print(f"[state] round trip K={len(pw)}", flush=True)
    st = engine.State(sy.spins, pw)
    try:
        st.set(sy.op("Ix_0+Iy_2"))
        got = st.fetch()
        assert np.array_equal(got, np.broadcast_to(sy.op("Ix_0+Iy_2"), got.shape))
        rho = rand_herm_stack(np.random.default_rng(3), len(pw), sy.N)
        st.load(rho)
        assert np.array_equal(st.fetch(), rho)
    finally:
        st.close()


@pytest.mark.parametrize("wr,repeat", [(10000.0, 1), (10000.0, 5), (0.0, 2)])
def test_apply_matches_oracle(wr, repeat):
    sy, pw, dts, t1s, rf = case3()
    rho = rand_herm_stack(np.random.default_rng(11), len(pw), sy.N)
    print(f"[state] apply wr={wr} repeat={repeat} K={len(pw)}", flush=True)
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        st.apply(sy, wr, MAGIC, dts, rf=rf, step_t1=t1s, repeat=repeat)
        got = st.fetch()
    finally:
        st.close()
    want = O.state_apply(sy, wr, MAGIC, dts, pw, rho, rf=rf, step_t1=t1s, repeat=repeat)
    err = rel(got, want)
    print(f"[state] apply wr={wr} repeat={repeat}: rel {err:.3e}", flush=True)
    assert err < 1e-12


def test_apply_in_batches_is_bitwise_the_same():
    """A memory budget that forces several crystallite batches changes nothing (5 spins, one
    32-state block under the rf: 40 steps keep about 0.7 MB per crystallite, the exp scratch
    is 4096 slots of 74 KB)."""
    sy = workloads.small_cluster(5, 77)
    pw = zcw(1154)
    dts = np.full(40, 0.2e-6)
    rf = [("1H", np.full(40, 80000.0), np.linspace(0, 1, 40), np.zeros(40))]
    out = []
    for budget in (0.0, 6.5e8):
        print(f"[state] batches budget={budget}", flush=True)
        st = engine.State(sy.spins, pw)
        try:
            st.set_memory_budget(budget)
            st.set(sy.op("Iz"))
            st.apply(sy, 15000.0, MAGIC, dts, rf=rf, repeat=3)
            out.append(st.fetch())
        finally:
            st.close()
    assert np.array_equal(out[0], out[1])
    want = O.state_apply(sy, 15000.0, MAGIC, dts, pw, np.stack([sy.op("Iz")] * len(pw)), rf=rf,
                         repeat=3)
    assert rel(out[0], want) < 1e-12


def test_trace_matches_numpy():
    sy, pw, *_ = case3()
    rho = rand_herm_stack(np.random.default_rng(4), len(pw), sy.N)
    det = sy.op("Ip_0+Ip_1")
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        got = st.trace(det)
        raw = st.trace(det, normalize=False)
    finally:
        st.close()
    tr = np.einsum("kab,ba->k", rho, det)
    want = np.sum(pw.weight * tr)
    assert abs(raw - want) < 1e-12 * abs(want)
    assert abs(got - want / np.trace(det @ det.conj().T)) < 1e-12 * abs(got)


def project_numpy(rho, op):
    """cycler(op): rho_k <- Tr(rho_k op^H) op / Tr(op op^H) (subSequence::propogate,
    sequencerun.cc:164-178), written with numpy's matrix trace."""
    opH = op.conj().T
    t = np.einsum("kab,ba->k", rho, opH) / np.trace(op @ opH)
    return t[:, None, None] * op[None]


@pytest.mark.parametrize("K,expr", [(51, "Ip_0+Iz_2"), (5000, "Ix_1")])
def test_project_matches_numpy(K, expr):
    """A non-Hermitian operator checks the conjugation; K=5000 exceeds the launch's 4096
    workgroups, so the crystallite loop runs twice on some of them."""
    sy, pw, *_ = case3()
    rng = np.random.default_rng(11)
    if K != len(pw):
        pw = Powder(rng.uniform(0, np.pi, K), rng.uniform(0, 2 * np.pi, K),
                    rng.uniform(0, 2 * np.pi, K), np.full(K, 1.0 / K), "random")
    rho = rand_herm_stack(rng, K, sy.N)
    op = sy.op(expr)
    print(f"[state] project K={K} op={expr}", flush=True)
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        st.project(op)
        got = st.fetch()
        with pytest.raises(engine.EngineError):
            st.project(np.zeros_like(op))
    finally:
        st.close()
    want = project_numpy(rho, op)
    assert rel(got, want) < 1e-13
    # a projection is idempotent: the projected state is a multiple of op
    assert rel(project_numpy(got, op), got) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("c64", 1e-5)])
def test_gamma_compute_fid_from_the_state(precision, tol):
    """Random Hermitian rho per crystallite: every block pair D couples is populated."""
    sy, pw, *_ = case3()
    rho = rand_herm_stack(np.random.default_rng(8), len(pw), sy.N)
    det = sy.op("Ip_2")
    print(f"[state] gamma-COMPUTE {precision} K={len(pw)}", flush=True)
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        g, swg = engine.state_fid(sy, 10000.0, MAGIC, 320000.0, 64, pw, st, det, t0=7e-6,
                                  precision=precision)
    finally:
        st.close()
    c, swc = O.state_fid(sy, 10000.0, MAGIC, 320000.0, 64, pw, rho, det, t0=7e-6)
    err = rel(g, c)
    print(f"[state] gamma-COMPUTE {precision}: rel {err:.3e}", flush=True)
    assert swg == swc and err < tol


def test_gamma_compute_fid_after_apply_larger_cluster():
    """A 5-spin cluster: set, apply a pulse, FID from the state; equals the oracle's prepared
    FID and the engine's own preparation stage."""
    sy = workloads.small_cluster(5, 77)
    pw = zcw(1154).subset(np.arange(0, 1154, 61))
    dts = np.array([0.6e-6, 0.6e-6, 0.6e-6, 2.0e-6])
    amp = np.array([90000.0, 90000.0, 90000.0, 0.0])
    rf = [("1H", amp, np.full(4, np.deg2rad(30.0)), np.array([2000.0, 2000.0, 2000.0, 0.0]))]
    rho, det = sy.op("Iz"), sy.op("Ip")
    wr, sw, npts, t0 = 15000.0, 15000.0 * 32, 96, float(dts.sum())
    print(f"[state] cluster5 apply + fid K={len(pw)}", flush=True)
    st = engine.State(sy.spins, pw)
    try:
        st.set(rho)
        st.apply(sy, wr, MAGIC, dts, rf=rf)
        g, _ = engine.state_fid(sy, wr, MAGIC, sw, npts, pw, st, det, t0=t0)
    finally:
        st.close()
    c, _ = O.prep_fid(sy, wr, MAGIC, sw, npts, pw, rho, det, dts, rf=rf, t0=t0)
    p, _ = engine.prep_fid(sy, wr, MAGIC, sw, npts, pw, rho, det, dts, rf=rf)
    print(f"[state] cluster5: rel vs oracle {rel(g, c):.3e}, vs prep {rel(g, p):.3e}", flush=True)
    assert rel(g, c) < 1e-10 and rel(g, p) < 1e-12


def test_static_fid_from_the_state():
    sy, pw, *_ = case3()
    rho = rand_herm_stack(np.random.default_rng(21), len(pw), sy.N)
    det = sy.op("Ip_0")
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        g, _ = engine.state_fid(sy, 0.0, 0.0, 40000.0, 128, pw, st, det)
    finally:
        st.close()
    c, _ = O.state_fid(sy, 0.0, 0.0, 40000.0, 128, pw, rho, det)
    print(f"[state] static: rel {rel(g, c):.3e}", flush=True)
    assert rel(g, c) < 1e-10


def test_pulse_fid_from_the_state():
    sy, pw, dts, t1s, rf = case3()
    rho = rand_herm_stack(np.random.default_rng(13), len(pw), sy.N)
    det = sy.op("Ip_0+Ip_1")
    st = engine.State(sy.spins, pw)
    try:
        st.load(rho)
        g = engine.state_pulse_fid(sy, 12000.0, MAGIC, dts, 40, pw, st, det, rf=rf, step_t1=t1s)
    finally:
        st.close()
    c = O.state_pulse_fid(sy, 12000.0, MAGIC, dts, 40, pw, rho, det, rf=rf, step_t1=t1s)
    print(f"[state] pulse fid: rel {rel(g, c):.3e}", flush=True)
    assert rel(g, c) < 1e-10


def test_state_size_mismatch_is_rejected():
    sy, pw, *_ = case3()
    st = engine.State(sy.spins, pw)
    p = engine.Plan(sy, 10000.0, MAGIC, 320000.0, 16)
    try:
        p.set_state(st)
        with pytest.raises(engine.EngineError) as e:
            p.powder_fid(pw.subset(np.arange(3)), sy.op("Ip"), sy.op("Ip"))
        assert e.value.status == 1
        other = engine.State(2, pw)
        try:
            with pytest.raises(engine.EngineError):
                p.set_state(other)
        finally:
            other.close()
    finally:
        p.close()
        st.close()
