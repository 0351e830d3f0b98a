"""Command line of the engine for solid-2.0 input files, the counterpart of ``solid file.sim``
(examples/solid-2.0/src/solid.cc:72-208):

    python -m blochlib_amd.solid file.sim [--precision fp64|c64] [--outdir DIR]

Reads the file, runs its pulses{} program on the GPU (the supported subset is described in
blochlib_amd/simfile.py) and writes the savefid outputs into DIR (default: the current
directory, as the reference does).
"""
from __future__ import annotations

import argparse
import sys

from .simfile import SimFile
from .spinsys import BlochLibError


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m blochlib_amd.solid")
    ap.add_argument("simfile")
    ap.add_argument("--precision", choices=["fp64", "c64", "bf16x3"], default="fp64")
    ap.add_argument("--outdir", default=".")
    args = ap.parse_args(argv)
    try:
        sim = SimFile.read(args.simfile)
        res = sim.run(outdir=args.outdir, precision=args.precision, verbose=True)
    except BlochLibError as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1
    for f in res["files"]:
        print(f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
