"""Powder orientation sets: the file reader of BlochLib's ``powder`` class.

``powder::read`` (src/QMspins/driver/powder.cc:333-389): columns are
``phi theta [gamma] weight``; 1 column -> theta 0, weight 1; 2 columns -> weight 1;
3 columns -> (phi, theta, weight) with gamma 0; 4+ columns -> (phi, theta, gamma, weight).
Lines starting with ``#`` or ``%`` are skipped.  Angles are radians.  Weights are NOT
renormalised by the reader (only the built-in generators normalise, powder.cc:310).

The ZCW_3_1154, ZCW_3_3722 and rep2000 sets shipped in ``crystals/`` are the reference's data
files (crystals/README).
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
    def single(cls, theta=0.0, phi=0.0, gamma=0.0, weight=1.0) -> "Powder":
        a = lambda v: np.array([v], dtype=np.float64)
        return cls(a(theta), a(phi), a(gamma), a(weight), "single")


def zcw(K: int) -> Powder:
    """The shipped ZCW 3-angle sets by size (1154 or 3722)."""
    names = {1154: "ZCW_3_1154", 3722: "ZCW_3_3722"}
    if K not in names:
        raise BlochLibError(f"no shipped ZCW set with {K} orientations")
    return Powder.read(os.path.join(CRYSTALS, names[K]))
