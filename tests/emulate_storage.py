"""Storage-format emulation on the oracle at the 4096x2048 precision
workload's size (shot 0, FP64 arithmetic): every array a half-step updates
is rounded to the format after that half-step, and the seismograms are
compared with the unrounded FP64 run.

Formats: "f16" and "bf16" with one power-of-two scale per array (the
library's FP16/BF16 storage), "blk16" block-scaled int16 (one power-of-two
quantum per 64 consecutive values).  Prints one JSON line per format with
the relative L2 error of the seismograms after 1000, 2000, 4000 steps and
the full run.  The full 6000 steps take about 30 minutes per format on 8
cores; this is an analysis script, not a test.

    python tests/emulate_storage.py [steps] [fmt ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from fd23_amd import _lib, workloads as W  # noqa: E402
from helpers import unpack_cpml  # noqa: E402

STRESS = ("U_x", "W_x", "U_z", "W_z")      # psi the stress half updates
VELOCITY = ("Txx_x", "Txz_x", "Txz_z", "Tzz_z")


def run(wl, nt, fmt, log_every=500):
    rx, rz = wl.receivers()
    stf, M = wl.stf.astype(np.float64), wl.M.astype(np.float64)
    xs, zs = wl.sources[0]
    cp = unpack_cpml(wl.packed, wl.wave, wl.nx, wl.nz, np.float64)
    sim = O.PSV(wl.nx, wl.nz, wl.npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), wl.mat)), cp, False,
                np.float64)
    su, sw = [], []
    t0 = time.time()
    for it in range(nt):
        sim.stress(it)
        sim.inject(0, it, stf, M, xs, zs)
        if fmt:
            for a in [sim.Txx, sim.Tzz, sim.Txz] + [sim.psi[n] for n in STRESS]:
                O.round_storage(a, fmt)
        sim.velocity(it)
        if fmt:
            for a in [sim.U, sim.W] + [sim.psi[n] for n in VELOCITY]:
                O.round_storage(a, fmt)
        su.append(sim.record("U", rx, rz))
        sw.append(sim.record("W", rx, rz))
        if log_every and it % log_every == 0:
            print(fmt or "fp64", it, round(time.time() - t0, 1), file=sys.stderr, flush=True)
    return np.stack([np.array(su), np.array(sw)])


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 6000
    fmts = sys.argv[2:] or ["f16", "bf16", "blk16"]
    wl = W.config_psv_precision(_lib.STORE_F64).build()
    ref = run(wl, nt, None)
    marks = sorted({k for k in (1000, 2000, 4000) if k < nt} | {nt})
    for fmt in fmts:
        e = run(wl, nt, fmt)
        out = {"format": fmt, "steps": nt, "shot": 0,
               "rel_l2": {k: float(np.linalg.norm(e[:, :k] - ref[:, :k]) / np.linalg.norm(ref[:, :k]))
                          for k in marks}}
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
