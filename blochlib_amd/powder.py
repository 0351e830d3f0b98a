"""Powder orientation sets: the file reader of BlochLib's ``powder`` class.

``powder::read`` (src/QMspins/driver/powder.cc:333-389): columns are
``phi theta [gamma] weight``; 1 column -> theta 0, weight 1; 2 columns -> weight 1;
3 columns -> (phi, theta, weight) with gamma 0; 4+ columns -> (phi, theta, gamma, weight).
Lines starting with ``#`` or ``%`` are skipped.  Angles are radians.  Weights are NOT
renormalised by the reader (only the built-in generators normalise, powder.cc:310).

The ZCW_3_1154, ZCW_3_3722 and rep2000 sets shipped in ``crystals/`` are the reference's data
files (crystals/README).  ``Powder.generate`` restates the built-in generators.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .spinsys import BlochLibError

CRYSTALS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crystals")


@dataclass
class Powder:
    theta: np.ndarray
    phi: np.ndarray
    gamma: np.ndarray
    weight: np.ndarray
    name: str = ""

    def __len__(self):
        return int(self.theta.shape[0])

    def subset(self, idx) -> "Powder":
        idx = np.asarray(idx)
        return Powder(self.theta[idx].copy(), self.phi[idx].copy(), self.gamma[idx].copy(),
                      self.weight[idx].copy(), self.name + "[subset]")

    @classmethod
    def read(cls, path: str) -> "Powder":
        if not os.path.exists(path):
            alt = os.path.join(CRYSTALS, os.path.basename(path))
            if os.path.exists(alt):
                path = alt
            else:
                raise BlochLibError(f"no POWDER file name like {path}")
        th, ph, ga, w = [], [], [], []
        with open(path, "r") as f:
            for line in f:
                tok = line.split()
                if not tok or tok[0][0] in "#%":
                    continue
                ph.append(float(tok[0]))
                th.append(float(tok[1]) if len(tok) >= 2 else 0.0)
                if len(tok) == 3:
                    w.append(float(tok[2]))
                    ga.append(0.0)
                elif len(tok) >= 4:
                    ga.append(float(tok[2]))
                    w.append(float(tok[3]))
                else:
                    w.append(1.0)
                    ga.append(0.0)
        f64 = np.float64
        return cls(np.array(th, f64), np.array(ph, f64), np.array(ga, f64), np.array(w, f64),
                   os.path.basename(path))

    @classmethod
    def generate(cls, method: str, tstep: int, pstep: int, gstep: int = 1) -> "Powder":
        """Built-in orientation sets of powder::calcPow (powder.cc:179-311): ``zcw`` (also
        ``zcw2d``), ``alderman``, ``sphere``, ``rect``, ``sophie``, ``liquid``; each set is
        repeated for max(gstep, 1) gamma values g/gstep * pi and the weights are normalised to
        sum 1.  Any other name is read as a file (powder.cc:98-111)."""
        m = {"zcw2d": "zcw", "zcw2D": "zcw"}.get(method, method)
        if m not in ("zcw", "alderman", "sphere", "rect", "sophie", "liquid"):
            return cls.read(method)
        th, ph, w = [], [], []
        pi = np.pi
        if m == "zcw":
            ts, ps = (21, 34, 55, 89, 144, 233, 377, 616, 987, 4184), (8, 13, 21, 34, 55, 89, 144,
                                                                         233, 377, 2584)
            if tstep not in ts or pstep not in ps or tstep <= pstep:
                raise BlochLibError("For using the ZCW method 'theta_step' and 'phi_step' must be "
                                    "fibonacci numbers where 'theta_step' > 'phi_step'")
            for k in range(1, tstep):
                fac = k / tstep
                f0 = fac - np.floor(fac)
                f1 = fac * pstep
                f1 = f1 - np.floor(f1)
                ph.append(2.0 * pi * f0)
                th.append(pi * f1)
                w.append(np.sin(pi * f1) / tstep)
        elif m == "sphere":
            for k in range(1, tstep + 1):
                for q in range(1, pstep + 1):
                    t = (pi / 2.0) * (k / tstep)
                    th.append(t)
                    ph.append(2 * pi * (q / (pstep + 1)))
                    w.append(np.sin(t))
        elif m == "rect":
            t = pi / (2.0 * tstep)
            while t <= pi / 2.0:
                nmax = 2.0 * np.sin(t) * pstep
                q = 0
                while q < nmax:
                    ph.append(2 * pi * q / nmax)
                    th.append(t)
                    w.append(np.sin(t))
                    q += 1
                t += pi / (2.0 * tstep)
        elif m == "sophie":
            for k in range(1, tstep):
                for q in range(1, k + 1):
                    th.append((pi / 2) * (k / (tstep - 1)))
                    ph.append((pi / 2) * ((k - q + 1) / k))
                    w.append(1.0)
        elif m == "alderman":
            for kk in range(2):
                for i in range(1, tstep + 1):
                    for j in range(1, tstep - i + 1):
                        R = np.sqrt(float(i * i + j * j + (tstep - i - j) ** 2))
                        w.append(1.0 / R ** 3)
                        th.append(np.arccos((tstep - i - j) / R))
                        ph.append(np.arctan(j / i) + kk * pi / 2)
        else:  # liquid
            th, ph, w = [0.0], [0.0], [1.0]
        ng = max(gstep, 1)
        gam = [0.0 if gstep == 0 else g / gstep * pi for g in range(ng)]
        f64 = np.float64
        T = np.tile(np.array(th, f64), ng)
        P = np.tile(np.array(ph, f64), ng)
        G = np.repeat(np.array(gam, f64), len(th)) if m != "liquid" else np.zeros(ng)
        W = np.tile(np.array(w, f64), ng)
        return cls(T, P, G, W / W.sum(), f"{method}({tstep},{pstep},{gstep})")

    @classmethod
    def single(cls, theta=0.0, phi=0.0, gamma=0.0, weight=1.0) -> "Powder":
        a = lambda v: np.array([v], dtype=np.float64)
        return cls(a(theta), a(phi), a(gamma), a(weight), "single")


def zcw(K: int) -> Powder:
    """The shipped ZCW 3-angle sets by size (1154 or 3722)."""
    names = {1154: "ZCW_3_1154", 3722: "ZCW_3_3722"}
    if K not in names:
        raise BlochLibError(f"no shipped ZCW set with {K} orientations")
    return Powder.read(os.path.join(CRYSTALS, names[K]))
