"""Writes tests/golden/config3_oracle_fid.npz: the FP64 CPU oracle's powder FID of the dense
precision configuration (10-spin 1H cluster, N = 1024, MAS 20 kHz, magic angle, 64 steps per
period, 512 points, rho = Ix, detect = I+, normalised by Tr(D D^H)) summed over four ZCW 1154
crystallites (indices 0, 300, 600, 900 with their file weights).

The oracle is the dense restatement of the reference algorithm (oracle/blfid_oracle.c), threaded
over the crystallites; at N = 1024 it needs ~4.4 GB per thread and tens of minutes, so the result
is kept as a fixture for the GPU test rather than recomputed there.

    python tests/golden/make_config3_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from blochlib_amd import workloads  # noqa: E402
from oracle import oracle as O  # noqa: E402

IDX = np.array([0, 300, 600, 900])


def main():
    w = workloads.config3()
    rho, det = w.operators()
    pw = w.powder().subset(IDX)
    t = time.time()
    fid, sw = O.powder_fid(w.system, w.wr, w.rotor, w.sw, w.npts, pw, rho, det,
                           nthreads=len(IDX))
    print(f"oracle: {len(IDX)} crystallites of {w.name} in {time.time() - t:.0f} s", flush=True)
    np.savez(os.path.join(HERE, "config3_oracle_fid.npz"), idx=IDX, fid=fid,
             sw=np.array([sw]), system=np.array(w.system.to_text()))


if __name__ == "__main__":
    main()
