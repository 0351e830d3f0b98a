"""GPU parity: the HIP engine (through the C ABI) against the FP64 CPU oracle.

Tolerance: relative L2 <= 1e-9 for the FP64 engine on these sizes (the engine's exp is a
Taylor/scaling-squaring evaluation and its eigen-solver is joint Jacobi, so agreement is to
rounding, not bitwise).  The accuracy bar of the workload is 1e-5.
"""
import sys

import numpy as np
import pytest

from blochlib_amd import engine, workloads
from blochlib_amd.powder import Powder, zcw
from blochlib_amd.spinsys import SolidSys
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-9
MAGIC = workloads.MAGIC_DEG


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def compare(name, sy, wr, rotor, sw, npts, pw, rho, det, t0=0.0, normalize=True, tol=TOL):
    print(f"[parity] {name}: S={sy.spins} K={len(pw)} wr={wr} sw={sw} npts={npts}", flush=True)
    g, swg = engine.powder_fid(sy, wr, rotor, sw, npts, pw, rho, det, t0=t0, normalize=normalize)
    c, swc = O.powder_fid(sy, wr, rotor, sw, npts, pw, rho, det, t0=t0, normalize=normalize)
    assert swg == swc
    err = rel_l2(g, c)
    print(f"[parity] {name}: rel-L2 {err:.3e}", flush=True)
    assert err < tol, (name, err)
    return g, c


SYS3 = """numspin 3
T 1H 0
T 1H 1
T 13C 2
C 1000 5000 0.5 0 10 20 30
C -1000 3000 0.2 1 40 50 60
C 300 7000 0.9 2
D -20000 0 1 0 30 40
D 5000 1 2 10 70 20
J 140 0 2
J 50 0 1 1
"""


def test_config1_full_powder():
    """Config 1 as specified: two dipolar-coupled 1H, MAS 10 kHz, 64 steps, all 1154 ZCW
    orientations, 256 points, in FP64 (1e-9) and in the complex64 form (1e-5)."""
    w = workloads.config1()
    pw = w.powder()
    rho, det = w.operators()
    g, c = compare("config1-full", w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det)
    g64, _ = engine.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det,
                               precision="c64")
    err = rel_l2(g64, c)
    print(f"[parity] config1-full complex64: rel-L2 {err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("sw", [80000.0, 70000.0])
@pytest.mark.parametrize("rho_name", ["Ix", "Iz+0.3*Ix_0"])
def test_hetero_strongJ_rosym_and_general_rho(sw, rho_name):
    sy = SolidSys.from_text(SYS3)
    pw = zcw(1154).subset(np.arange(0, 1154, 97))
    compare(f"sys3-{rho_name}-sw{sw}", sy, 10000.0, MAGIC, sw, 40, pw, sy.op(rho_name),
            sy.op("Ip"), normalize=False)


def test_start_time_and_other_rotor_angle():
    sy = SolidSys.from_text(SYS3)
    pw = zcw(1154).subset(np.arange(5, 1154, 131))
    compare("sys3-t0", sy, 7000.0, 30.0, 7000.0 * 12, 50, pw, sy.op("Ix"), sy.op("Ip"), t0=3.1e-5)


@pytest.mark.parametrize("S", [4, 5, 6])
def test_clusters(S):
    sy = workloads.small_cluster(S, 77 + S)
    pw = zcw(1154).subset(np.arange(0, 1154, 191))
    compare(f"cluster{S}", sy, 15000.0, MAGIC, 15000.0 * 16, 64, pw, sy.op("Ix"), sy.op("Ip"))


@pytest.mark.parametrize("S", [7, 8])
def test_large_blocks(S):
    sy = workloads.small_cluster(S, 300 + S)
    pw = zcw(1154).subset([3, 400])
    compare(f"cluster{S}", sy, 15000.0, MAGIC, 15000.0 * 8, 24, pw, sy.op("Ix"), sy.op("Ip"))


def test_static_path_matches_reference_output():
    sy = SolidSys.from_text("numspin 2\nT 1H 0\nT 1H 1\nC 5000  2134 0 0\nC -5000 2789 0.5 1\n")
    g, c = compare("static-simpSTA", sy, 0.0, 0.0, 40000.0, 512, Powder.read("rep2000"),
                   sy.op("Ix"), sy.op("Ip"))
    ref = np.loadtxt("tests/golden/simpSTA.txt")
    fr = ref[:, 4] + 1j * ref[:, 5]
    assert np.abs(g - fr).max() < 2e-6


def test_static_dipolar_rotor_tilted():
    sy = SolidSys.from_text(SYS3)
    pw = zcw(1154).subset(np.arange(0, 1154, 61))
    compare("static-sys3-rotor30", sy, 0.0, 30.0, 50000.0, 64, pw, sy.op("Ix"), sy.op("Ip"))


def test_empty_powder_and_single_point():
    sy = SolidSys.from_text(SYS3)
    p = engine.Plan(sy, 10000.0, MAGIC, 80000.0, 1)
    f = p.powder_fid(Powder(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0)), sy.op("Ix"),
                     sy.op("Ip"))
    assert f.shape == (1,) and f[0] == 0
    g, c = compare("npts1", sy, 10000.0, MAGIC, 80000.0, 1, zcw(1154).subset([0, 1, 2]),
                   sy.op("Ix"), sy.op("Ip"))
    assert np.isclose(g[0], 2.0 * 3 / 1154)


def test_repeat_runs_are_consistent_and_batches_agree():
    w = workloads.config1()
    pw = w.powder().subset(np.arange(0, 1154, 7))
    rho, det = w.operators()
    p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts)
    a = p.powder_fid(pw, rho, det)
    st = p.stats()
    p.set_memory_budget(st["fixed_bytes"] + 3.5 * st["bytes_per_crystallite"])  # batches of 3
    p.run(True)
    b = p.fetch()
    # the powder sum runs in a fixed order (crystallite groups of 64 carried across batches),
    # so the batch size does not change a single bit
    assert np.array_equal(b, a)


def test_shared_workspace_two_plans_small_batches():
    """Two plans borrowing one workspace small enough to force several crystallite batches."""
    w1 = workloads.config1()
    sy2 = workloads.small_cluster(5, 5)
    pw = zcw(1154).subset(np.arange(0, 1154, 53))
    p1 = engine.Plan(w1.system, w1.wr, w1.rotor, w1.sw, w1.npts)
    p2 = engine.Plan(sy2, 15000.0, MAGIC, 15000.0 * 16, 64)
    need = max(p1.stats()["bytes_per_crystallite"], p2.stats()["bytes_per_crystallite"])
    fixed = max(p1.stats()["fixed_bytes"], p2.stats()["fixed_bytes"])
    ws = engine.Workspace(fixed + 4.5 * need)
    for p, sy, (wr, sw, npts) in [(p1, w1.system, (w1.wr, w1.sw, w1.npts)),
                                  (p2, sy2, (15000.0, 15000.0 * 16, 64))]:
        p.set_workspace(ws)
        got = p.powder_fid(pw, sy.op("Ix"), sy.op("Ip"))
        ref, _ = O.powder_fid(sy, wr, MAGIC if p is p2 else w1.rotor, sw, npts, pw, sy.op("Ix"),
                              sy.op("Ip"))
        print(f"[parity] workspace plan S={sy.spins}: rel-L2 {rel_l2(got, ref):.3e}", flush=True)
        assert rel_l2(got, ref) < TOL
    p1.close()
    p2.close()
    ws.close()


def test_blocks_beyond_lds_nine_spins():
    """9 spins: magnetisation blocks of 126 states, beyond the LDS eigen-solver tile (80)."""
    sy = workloads.small_cluster(9, 409)
    pw = zcw(1154).subset([7])
    compare("cluster9", sy, 20000.0, MAGIC, 20000.0 * 4, 8, pw, sy.op("Ix"), sy.op("Ip"))


def test_ten_spin_dense_shape_invariants():
    """The 10-spin (N=1024, blocks up to 252) workload at its real step count and FID length
    on a few crystallites: fid[0] equals its closed form 2*sum(w) for rho = Fx, det = Fp, the
    signal stays bounded by it, and batching the crystallites does not change the result."""
    w = workloads.config3()
    pw = w.powder().subset([0, 300, 700, 1100])
    rho, det = w.operators()
    print("[parity] config3 shape invariants", flush=True)
    p = engine.Plan(w.system, w.wr, w.rotor, w.sw, w.npts)
    st = p.stats()
    assert st["max_block"] == 252
    a = p.powder_fid(pw, rho, det)
    assert np.isclose(a[0].real, 2.0 * pw.weight.sum(), rtol=1e-12) and abs(a[0].imag) < 1e-12
    assert np.all(np.isfinite(a)) and np.abs(a).max() <= abs(a[0]) * (1 + 1e-9)
    p.set_memory_budget(st["fixed_bytes"] + 2.5 * st["bytes_per_crystallite"])
    p.run(True)
    b = p.fetch()
    assert np.array_equal(b, a)
    p.close()


@pytest.mark.parametrize("det_name", ["Ix", "Im"])
def test_rosym_detector_variants(det_name):
    """rho = (det + det^H)/2 with a Hermitian detector (both block orders of det non-zero: R is
    formed by its own products) and with a lowering detector (R = Q^H elementwise)."""
    sy = workloads.small_cluster(5, 91)
    det = sy.op(det_name)
    rho = 0.5 * (det + det.conj().T)
    pw = zcw(1154).subset(np.arange(0, 1154, 151))
    compare(f"rosym-{det_name}", sy, 15000.0, MAGIC, 15000.0 * 16, 48, pw, rho, det,
            normalize=False)


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
@pytest.mark.parametrize("npts", [40, 300])
def test_correlation_fft_sizes(n, npts):
    """Every FFT size of the correlation (mixed-radix passes 8+2, 8+4, 8+8, 8+8+2, 8+8+4,
    8+8+8) with the FID sums held in registers (40 points) and, where the period table does not
    fit, accumulated per chunk (300 points); general rho (R from its own products) and rosym rho
    (R formed from Q), FP64 and complex64."""
    sy = SolidSys.from_text(SYS3)
    pw = zcw(1154).subset(np.array([5, 400, 777, 1100]))
    wr = 10000.0
    for rho_name in ("Iz+0.3*Ix_0", "Ix"):
        compare(f"fft-n{n}-p{npts}-{rho_name}", sy, wr, MAGIC, wr * n, npts, pw,
                sy.op(rho_name), sy.op("Ip"), normalize=False)
        print(f"[parity] fft-n{n}-p{npts}-{rho_name}: complex64", flush=True)
        g, _ = engine.powder_fid(sy, wr, MAGIC, wr * n, npts, pw, sy.op(rho_name), sy.op("Ip"),
                                 normalize=False, precision="c64")
        c, _ = O.powder_fid(sy, wr, MAGIC, wr * n, npts, pw, sy.op(rho_name), sy.op("Ip"),
                            normalize=False)
        err = rel_l2(g, c)
        print(f"[parity] fft-n{n}-p{npts}-{rho_name}: complex64 rel-L2 {err:.3e}", flush=True)
        assert err < 1e-5
