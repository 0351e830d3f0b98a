"""Seeded synthetic workloads (reference-syntax spin systems) for the five benchmark shapes.

Every generator emits BlochLib ``spins{}`` text (``T``, ``C``, ``D``, ``J`` lines) so the same
input feeds the GPU engine, the CPU oracle and BlochLib itself.  Seeds are numpy PCG64 with
seed = 1000 * config + index.

Dipolar couplings b = -120120 Hz A^3 / r^3 (1H-1H) and -30211 Hz A^3 / r^3 (13C-1H); the
dipolar PAS z axis (column 2 of rotationMatrix3D(alpha,beta,gamma), space_ten.cc:578-618, i.e.
(-sin b cos g, sin b sin g, cos b)) is put on the internuclear vector u with alpha = 0,
beta = acos(u_z), gamma = atan2(u_y, -u_x).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .powder import Powder, zcw
from .spinsys import SolidSys

MAGIC_DEG = 54.7356
B_HH = -120120.0
B_CH = -30211.0


@dataclass
class Workload:
    name: str
    system: SolidSys
    wr: float
    rotor: float
    sw: float
    npts: int
    powder_size: int
    rho: str = "Ix"
    det: str = "Ip"
    meta: dict = field(default_factory=dict)

    @property
    def steps(self) -> int:
        return int(np.floor(self.sw / self.wr + 0.5)) if self.wr else 1

    def powder(self) -> Powder:
        return zcw(self.powder_size)

    def operators(self):
        return self.system.op(self.rho), self.system.op(self.det)


def _euler_deg(rng):
    return [float(x) for x in rng.uniform(0.0, 360.0, size=3)]


def _dip_angles(u):
    u = u / np.linalg.norm(u)
    beta = np.degrees(np.arccos(np.clip(u[2], -1.0, 1.0)))
    gamma = np.degrees(np.arctan2(u[1], -u[0]))
    return 0.0, float(beta), float(gamma)


def config1(seed_index: int = 0) -> Workload:
    """Two dipolar-coupled 1H (N=4), MAS 10 kHz, 64 steps/period, ZCW 1154, 256 points."""
    rng = np.random.default_rng(1000 * 1 + seed_index)
    sy = SolidSys(isotopes=["1H", "1H"])
    sy.add_csa(1000.0, 5000.0, 0.5, 0, *_euler_deg(rng))
    sy.add_csa(-1000.0, 5000.0, 0.5, 1, *_euler_deg(rng))
    sy.add_dip(-20000.0, 0, 1, 0.0, 0.0, 0.0)
    return Workload("config1-plain-2x1H", sy, 10000.0, MAGIC_DEG, 640000.0, 256, 1154)


def cluster(nspins: int, rng, box=6.0, dmin=1.5, csa_range=(1000.0, 8000.0), iso=0.0):
    """Homonuclear 1H cluster: positions uniform in a box with rejection at r < dmin."""
    pos = []
    while len(pos) < nspins:
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - q) >= dmin for q in pos):
            pos.append(p)
    sy = SolidSys(isotopes=["1H"] * nspins)
    for i in range(nspins):
        sy.add_csa(iso, float(rng.uniform(*csa_range)), float(rng.uniform(0.0, 1.0)), i,
                   *_euler_deg(rng))
    for i in range(nspins):
        for j in range(i + 1, nspins):
            d = pos[j] - pos[i]
            r = np.linalg.norm(d)
            sy.add_dip(B_HH / r ** 3, i, j, *_dip_angles(d))
    return sy


def config2(nclusters: int = 64) -> list:
    """64 homonuclear 1H clusters of 4-8 spins, MAS 15 kHz, 128 steps/period, ZCW 3722,
    1024 points."""
    out = []
    for i in range(nclusters):
        rng = np.random.default_rng(1000 * 2 + i)
        S = int(rng.integers(4, 9))
        sy = cluster(S, rng)
        out.append(Workload(f"config2-cluster{i:02d}-S{S}", sy, 15000.0, MAGIC_DEG,
                            128 * 15000.0, 1024, 3722, meta={"spins": S}))
    return out


def config3() -> Workload:
    """10-spin 1H cluster (N=1024), MAS 20 kHz, 64 steps/period, ZCW 1154, 512 points."""
    rng = np.random.default_rng(1000 * 3)
    sy = cluster(10, rng)
    return Workload("config3-dense-10x1H", sy, 20000.0, MAGIC_DEG, 64 * 20000.0, 512, 1154)


def small_cluster(nspins: int, seed: int) -> SolidSys:
    return cluster(nspins, np.random.default_rng(seed))


@dataclass
class PulseWorkload:
    """A rotor-synchronised pulse program: one rotor period of `steps` equal steps with rf
    channels (isotope, amp_hz[steps], phase_rad[steps], offset_hz[steps]), one FID point per
    period (solid-2.0's ptop())."""
    name: str
    system: SolidSys
    wr: float
    rotor: float
    steps: int
    npts: int
    powder_size: int
    rf: list
    rho: str = "Ix_0"
    det: str = "Ip_0"
    meta: dict = field(default_factory=dict)

    def powder(self) -> Powder:
        return zcw(self.powder_size)

    def operators(self):
        return self.system.op(self.rho), self.system.op(self.det)


def phase_alternating_rf(isotope: str, steps: int, amp: float, cycles: int) -> list:
    """rf channel of constant amplitude whose phase alternates 0 / 180 degrees every half
    cycle, `cycles` full cycles per rotor period (steps must be a multiple of 2 * cycles)."""
    if steps % (2 * cycles):
        raise ValueError("steps per period must be a multiple of 2 * cycles")
    half = steps // (2 * cycles)
    ph = np.where((np.arange(steps) // half) % 2 == 0, 0.0, np.pi)
    return [(isotope, np.full(steps, float(amp)), ph, np.zeros(steps))]


def ch6_geometry(seed: int = 1000 * 4, r_ch: float = 1.09, dmin_hh: float = 1.0):
    """Six hydrogen positions at r_ch around a carbon at the origin, directions uniform on the
    sphere with rejection when two hydrogens are closer than dmin_hh."""
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < 6:
        u = rng.normal(size=3)
        p = r_ch * u / np.linalg.norm(u)
        if all(np.linalg.norm(p - q) >= dmin_hh for q in pos):
            pos.append(p)
    return np.array(pos)


def config4_system(set_index: int, geometry=None) -> SolidSys:
    """One 13C + six 1H (N=128) parameter set: seeded geometry shared by all sets, J(C-H) =
    140 Hz, seeded CSA per set (13C delta 2-12 kHz, 1H delta 1-8 kHz, eta uniform [0, 1],
    Euler angles uniform)."""
    H = ch6_geometry() if geometry is None else geometry
    rng = np.random.default_rng(1000 * 4 + 1 + set_index)
    sy = SolidSys(isotopes=["13C"] + ["1H"] * 6)
    sy.add_csa(0.0, float(rng.uniform(2000.0, 12000.0)), float(rng.uniform(0.0, 1.0)), 0,
               *_euler_deg(rng))
    for i in range(6):
        sy.add_csa(0.0, float(rng.uniform(1000.0, 8000.0)), float(rng.uniform(0.0, 1.0)), i + 1,
                   *_euler_deg(rng))
    for i in range(6):
        r = np.linalg.norm(H[i])
        sy.add_dip(B_CH / r ** 3, 0, i + 1, *_dip_angles(H[i]))
    for i in range(6):
        for j in range(i + 1, 6):
            d = H[j] - H[i]
            sy.add_dip(B_HH / np.linalg.norm(d) ** 3, i + 1, j + 1, *_dip_angles(d))
    for i in range(6):
        sy.add_j(140.0, 0, i + 1)
    return sy


def config4(nsets: int = 512) -> list:
    """13C + 6 1H under MAS 12.5 kHz with 100 kHz phase-alternating 1H rf (8 cycles per
    period), 256 steps per period, ZCW 1154, 2048 rotor-synchronised points, rho = Ix_0,
    det = Ip_0; one workload per CSA parameter set."""
    geo = ch6_geometry()
    steps = 256
    rf = phase_alternating_rf("1H", steps, 100000.0, 8)
    return [PulseWorkload(f"config4-CH6-set{i:03d}", config4_system(i, geo), 12500.0, MAGIC_DEG,
                          steps, 2048, 1154, rf, meta={"set": i})
            for i in range(nsets)]
