"""The oracle's box loop (PSV/psv_clean_double/main.cpp:153-200) against
what the physics fixes, on the CPU.

In a model made entirely of the box medium, the plane wave the box is
driven with is (up to the scheme's dispersion) an exact solution, so
- the residual grid, which carries the total field inside the box, equals
  the analytic incident wave at each field's time level: velocities at
  (nt+1/2) dt, stresses at nt dt after nt steps;
- almost nothing reaches the global grid outside the box.
A half-step timing slip, a wrong polarisation, a wrong stress factor or a
swapped ring (add/sub) each fails one of these.
"""
import numpy as np
import pytest

import oracle as O
from fd23_amd import host
from helpers import rel, unpack_cpml

NX, NZ, NPML, H, DT = 200, 150, 20, 10.0, 8e-4
VP, VS, RHO = 3000.0, 1732.0, 2300.0
X0, Z0, BX, BZ = 40, 40, 120, 70


def _run(incident, angle, nt, sfac=None):
    vp = np.full((NZ, NX), VP)
    vs = np.full((NZ, NX), VS)
    den = np.full((NZ, NX), RHO)
    mat, _ = host.psv_material(vp, vs, den, H, DT, f64=True)
    cp = unpack_cpml(host.cpml_pack(0, NX, NZ, NPML, DT, 1e-6, VP / H, 15.0), 0, NX, NZ, np.float64)
    sim = O.PSV(NX, NZ, NPML, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), mat)), cp, dtype=np.float64)
    box = O.PlaneWaveBox(sim, X0, Z0, BX, BZ, incident, angle, 15.0, 0.1, 1.0, VP, VS, RHO, H, DT)
    if sfac:
        box.sfac.update(sfac)
    O.run_psv_box(sim, box, nt)
    return sim, box


def _errors(box, nt, shift=0.0):
    inner = (slice(10, -10), slice(10, -10))
    out = {}
    for name, lev in (("U", 0.5), ("W", 0.5), ("Txx", 0.0), ("Tzz", 0.0), ("Txz", 0.0)):
        out[name] = rel(box.r[name][inner], -box.drive(name, (nt + lev + shift) * DT)[inner])
    return out


def _leak(sim, box):
    outside = np.abs(sim.U).copy()
    outside[Z0:Z0 + BZ, X0:X0 + BX] = 0
    return outside.max() / np.abs(box.r["U"]).max()


@pytest.mark.parametrize("incident,angle", [(0, 20.0), (1, -25.0)])
def test_box_carries_plane_wave(incident, angle):
    nt = 220
    sim, box = _run(incident, angle, nt)
    errs = _errors(box, nt)
    assert max(errs.values()) < 1e-2, errs
    # a half-step timing slip is an order of magnitude worse
    assert min(_errors(box, nt, 0.5).values()) > 3e-2
    assert _leak(sim, box) < 5e-3


def test_box_wrong_stress_factor_leaks():
    """A planted sign error in the shear-stress factor of the drive makes the
    box radiate and the residual drift from the incident wave."""
    nt = 220
    good_sim, good = _run(0, 20.0, nt)
    th = np.deg2rad(20.0)
    mu = RHO * VS * VS
    bad_sim, bad = _run(0, 20.0, nt, sfac={"Txz": +(2 * mu * np.sin(th) * -np.cos(th)) / VP})
    assert _leak(bad_sim, bad) > 10 * _leak(good_sim, good)
