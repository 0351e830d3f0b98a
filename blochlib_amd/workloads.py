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
