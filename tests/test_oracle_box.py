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
from helpers import point_source_frames, rel, unpack_cpml

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


# ---------------------------------------------------------------------------
# the reference's own box, driven by displacement frames
# (PSV/psv_clean_double/box/box.cpp:10-210, material/material.cpp:39-99)
# ---------------------------------------------------------------------------
FB = dict(xs=100, zs=75, bw=32)


def _homogeneous_sim(dtype=np.float64):
    vp = np.full((NZ, NX), VP)
    vs = np.full((NZ, NX), VS)
    den = np.full((NZ, NX), RHO)
    f64 = dtype == np.float64
    mat, _ = host.psv_material(vp, vs, den, H, DT, f64=f64)
    cp = unpack_cpml(host.cpml_pack(0, NX, NZ, NPML, DT, 1e-6, VP / H, 15.0), 0, NX, NZ, dtype)
    return O.PSV(NX, NZ, NPML, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), mat)), cp, dtype=dtype)


def box_frames(nt, xs, zs, bw, f0=15.0):
    """Frames of an explosive point source at (xs, zs) in the homogeneous
    model (helpers.point_source_frames) and that run's final U field."""
    sim = _homogeneous_sim()
    sim.dt_frames = DT
    stf = host.ricker(nt, DT, f0)
    M = host.psv_moment(0, 0, 0, 0, 0, H, DT, explosion=True).astype(np.float64)
    frames = point_source_frames(sim, nt, xs, zs, bw, stf, M)
    return frames, sim.U.copy()


def _file_box_run(frames, nt, srcpvel=VP, srcsvel=VS, srcden=RHO, dtype=np.float64):
    sim = _homogeneous_sim(dtype)
    box = O.FileBox(sim, FB["xs"], FB["zs"], FB["bw"], frames, srcpvel, srcsvel, srcden, H, DT)
    O.run_psv_box(sim, box, nt)
    return sim, box


def _outside(a, box, margin=0):
    m = np.ones(a.shape, bool)
    m[box.z0 + margin:box.z0 + box.bz - margin, box.x0 + margin:box.x0 + box.bx - margin] = False
    return a[m]


def test_file_box_rings_and_material_rule():
    sim = _homogeneous_sim()
    box = O.FileBox(sim, 100, 75, 24, np.zeros((2, 2, 32, 32)), VP, VS, RHO, H, DT)
    # BOX::init_box's index lists for bw = 24 (box.cpp:58-95)
    assert box.sub.sum() == 320 and box.add.sum() == 192
    assert (box.x0, box.z0) == (100 - 11, 75 - 11)
    assert box.nforced == 0  # the host material of the same medium is within 1e-4
    # one cell off by 2e-4 relative is forced, one off by 5e-5 is not
    sim = _homogeneous_sim()
    sim.mat["MU"][box.z0 + 3, box.x0 + 5] *= 1 + 2e-4
    sim.mat["LAM"][box.z0 + 7, box.x0 + 2] *= 1 + 5e-5
    b2 = O.FileBox(sim, 100, 75, 24, np.zeros((2, 2, 32, 32)), VP, VS, RHO, H, DT)
    assert b2.nforced == 1
    mu = RHO * VS * VS * DT / H
    assert sim.mat["MU"][box.z0 + 3, box.x0 + 5] == mu and sim.mat["MUA"][box.z0 + 3, box.x0 + 5] == mu


def test_file_box_drive_of_linear_frames():
    """u = a x, w = b z on the frame grid: every order of the reduced
    stencil differentiates a line exactly, so the drive is the closed form
    Txx = (GAM a + LAM b)/dt, Tzz = (LAM a + GAM b)/dt, Txz = 0, and the
    velocities are the frame differences over dt."""
    bw = 32
    sim = _homogeneous_sim()
    z, x = np.mgrid[0:bw + 8, 0:bw + 8].astype(np.float64)
    a, b = 3e-3, -2e-3
    frames = np.stack([np.stack([0.5 * a * x, 0.5 * b * z]), np.stack([a * x, b * z])])
    box = O.FileBox(sim, FB["xs"], FB["zs"], bw, frames, VP, VS, RHO, H, DT)
    s = box.source(0)
    lam, mu = box.bmat["LAM"], box.bmat["MU"]
    gam = lam + 2.0 * mu
    assert np.allclose(s["Txx"], (gam * a + lam * b) / DT, rtol=1e-12, atol=0)
    assert np.allclose(s["Tzz"], (lam * a + gam * b) / DT, rtol=1e-12, atol=0)
    assert np.abs(s["Txz"]).max() < 1e-12 * np.abs(s["Txx"]).max()
    assert np.allclose(s["U"], 0.5 * a * x[4:-4, 4:-4] / DT, rtol=1e-12)
    assert np.allclose(s["W"], 0.5 * b * z[4:-4, 4:-4] / DT, rtol=1e-12)


def test_file_box_reproduces_the_point_source():
    """Frames taken from a run of the same scheme make the box transparent:
    the residual grid (the scattered field) stays at rounding level, and
    outside the box the global grid equals the run that had the point
    source.  A one-step shift of the frames, or a source medium that
    differs from the model (forced into it), breaks both."""
    nt = 260
    frames, u_ref = box_frames(nt, FB["xs"], FB["zs"], FB["bw"])
    sim, box = _file_box_run(frames, nt)
    smax = max(np.abs(box.source(nt - 1)[k]).max() for k in ("U", "W"))
    leak = max(np.abs(box.r[k]).max() for k in ("U", "W")) / smax
    out = rel(_outside(sim.U, box), _outside(u_ref, box))
    print("file box leak", leak, "outside vs point source", out)
    assert leak < 1e-9 and out < 1e-9
    shifted = np.concatenate([frames[1:], frames[-1:]])
    sim2, box2 = _file_box_run(shifted, nt)
    leak2 = max(np.abs(box2.r[k]).max() for k in ("U", "W")) / smax
    assert leak2 > 1e-3 and rel(_outside(sim2.U, box2), _outside(u_ref, box2)) > 1e-3
    sim3, box3 = _file_box_run(frames, nt, srcpvel=VP * 1.01)
    assert box3.nforced == FB["bw"] ** 2
    assert rel(_outside(sim3.U, box3), _outside(u_ref, box3)) > 1e-3
