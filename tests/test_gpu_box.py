"""Hybrid box coupling on the GPU vs the oracle's box loop.

The oracle follows PSV/psv_clean_double/main.cpp:153-200 (ring add/sub of
box/box.cpp:198-210, residual passes of fd/cu_fd.cu:93-207) with the
plane-wave drive; both sides get the same heterogeneous model, so the
global grid carries a scattered field.  Tolerances: 1e-5 relative L2 for
FP32, 1e-10 for FP64, with field errors measured against the total-field
scale (inside the box the global grid holds total minus incident, a small
difference of large numbers).
"""
import numpy as np
import pytest

import oracle as O
from fd23_amd import _lib, host
from fd23_amd.core import Context
from helpers import layered_model, point_source_frames, rel, unpack_cpml

pytestmark = pytest.mark.gpu

NX, NZ, NPML, H, DT = 240, 180, 20, 10.0, 8e-4
BOX = dict(x0=60, z0=50, bx=120, bz=90)
MED = dict(vp=3000.0, vs=1732.0, rho=2300.0)


def _setup(f64, incident, angle):
    vp, vs, den = layered_model(NX, NZ, 77, vmin=2000.0, vmax=4500.0)
    if f64:
        vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
    mat, _ = host.psv_material(vp, vs, den, H, DT, f64=f64)
    packed = host.cpml_pack(_lib.WAVE_PSV, NX, NZ, NPML, DT, 1e-6, float(vp.max()) / H, 12.0)
    wave = dict(incident=incident, angle_deg=angle, f0=12.0, t0=1.5 / 12.0, amp=1.0)
    return mat, packed, wave


def _oracle(mat, packed, wave, f64, nt, rx, rz):
    dt_ = np.float64 if f64 else np.float32
    cp = unpack_cpml(packed, _lib.WAVE_PSV, NX, NZ, dt_)
    sim = O.PSV(NX, NZ, NPML, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in mat])), cp,
                dtype=dt_)
    box = O.PlaneWaveBox(sim, BOX["x0"], BOX["z0"], BOX["bx"], BOX["bz"], wave["incident"], wave["angle_deg"],
                         wave["f0"], wave["t0"], wave["amp"], MED["vp"], MED["vs"], MED["rho"], H, DT)
    u, w = O.run_psv_box(sim, box, nt, rx, rz)
    return sim, box, u, w


def _gpu(mat, packed, wave, f64, nt, rx, rz, staged):
    storage = _lib.STORE_F64 if f64 else _lib.STORE_F32
    ctx = Context(_lib.WAVE_PSV, NX, NZ, NPML, storage=storage, max_frames=nt, max_rec=rx.size)
    ctx.set_material(mat)
    ctx.set_cpml(packed)
    for c in range(2):
        ctx.set_receivers(0, c, rx, rz)
    ctx.box_configure(BOX["x0"], BOX["z0"], BOX["bx"], BOX["bz"], wave["incident"], wave["angle_deg"],
                      wave["f0"], wave["t0"], wave["amp"], MED["vp"], MED["vs"], MED["rho"], H, DT)
    if staged:
        for it in range(nt):
            ctx.step_stress()
            ctx.step_velocity()
            ctx.record(it)
    else:
        ctx.run(0, nt)
    ctx.sync()
    return ctx


@pytest.mark.parametrize("f64,incident,angle,staged", [
    (False, 0, 20.0, False),
    (False, 1, -15.0, False),
    (False, 0, 0.0, True),
    (True, 1, 25.0, False),
])
def test_box_matches_oracle(f64, incident, angle, staged):
    print("case: box f64=%s incident=%d angle=%g staged=%s" % (f64, incident, angle, staged), flush=True)
    nt = 300
    rx = np.arange(2, NX - 2, 3, dtype=np.int32)
    rz = np.full(rx.size, 25, np.int32)
    mat, packed, wave = _setup(f64, incident, angle)
    sim, box, u, w = _oracle(mat, packed, wave, f64, nt, rx, rz)
    tol = 1e-10 if f64 else 1e-5
    errs = {}
    with _gpu(mat, packed, wave, f64, nt, rx, rz, staged) as ctx:
        res = ctx.box_residual()
        # Errors are measured against the scale of the total field of the
        # same kind (velocities U, W; stresses Txx, Tzz, Txz): inside the box
        # the global grid holds total minus incident, a small difference,
        # and a component the incident wave lacks (U and Txz at vertical
        # incidence) is small in the residual grid too.
        kinds = {"U": ("U", "W"), "W": ("U", "W"), "Txx": ("Txx", "Tzz", "Txz"),
                 "Tzz": ("Txx", "Tzz", "Txz"), "Txz": ("Txx", "Tzz", "Txz")}
        for name in ("U", "W", "Txx", "Tzz", "Txz"):
            scale = max(np.linalg.norm(box.r[k]) for k in kinds[name])
            errs["resid " + name] = float(np.linalg.norm(res[name].astype(np.float64) - box.r[name]) / scale)
            g, o = ctx.field(name), getattr(sim, name)
            errs["field " + name] = float(np.linalg.norm(g.astype(np.float64) - o) / scale)
        errs["seis U"] = rel(ctx.seismograms(0, 0, nt), u)
        errs["seis W"] = rel(ctx.seismograms(0, 1, nt), w)
    print(errs, flush=True)
    assert np.abs(u).max() > 0
    bad = {k: v for k, v in errs.items() if not v < tol}
    assert not bad, bad


def test_box_homogeneous_plane_wave():
    """In the box medium everywhere, the residual grid carries the incident
    plane wave and almost nothing leaks into the global grid."""
    print("case: box homogeneous plane wave", flush=True)
    nt = 300
    vp = np.full((NZ, NX), MED["vp"], np.float32)
    vs = np.full((NZ, NX), MED["vs"], np.float32)
    den = np.full((NZ, NX), MED["rho"], np.float32)
    mat, _ = host.psv_material(vp, vs, den, H, DT)
    packed = host.cpml_pack(_lib.WAVE_PSV, NX, NZ, NPML, DT, 1e-6, MED["vp"] / H, 12.0)
    rx = np.array([10], np.int32)
    rz = np.array([10], np.int32)
    for incident in (0, 1):
        wave = dict(incident=incident, angle_deg=20.0, f0=12.0, t0=1.5 / 12.0, amp=1.0)
        with _gpu(mat, packed, wave, False, nt, rx, rz, False) as ctx:
            res = ctx.box_residual()
            peak = np.abs(res["U"]).max()
            x0, z0, bx, bz = BOX["x0"], BOX["z0"], BOX["bx"], BOX["bz"]
            outside = np.abs(ctx.field("U"))
            outside[z0:z0 + bz, x0:x0 + bx] = 0
            print("incident", incident, "peak", peak, "leak", outside.max() / peak, flush=True)
            assert peak > 0.1
            assert outside.max() / peak < 1e-2


def test_box_rejects_bad_geometry():
    with Context(_lib.WAVE_PSV, NX, NZ, NPML) as ctx:
        with pytest.raises(_lib.Fd23Error):
            ctx.box_configure(1, 50, 120, 90, 0, 0.0, 12.0, 0.1, 1.0, 3000.0, 1732.0, 2300.0, H, DT)
        with pytest.raises(_lib.Fd23Error):
            ctx.box_configure(60, 50, 8, 90, 0, 0.0, 12.0, 0.1, 1.0, 3000.0, 1732.0, 2300.0, H, DT)
    with Context(_lib.WAVE_SH, NX, NZ, NPML) as ctx:
        with pytest.raises(_lib.Fd23Error):
            ctx.box_configure(60, 50, 120, 90, 0, 0.0, 12.0, 0.1, 1.0, 3000.0, 1732.0, 2300.0, H, DT)


# ---------------------------------------------------------------------------
# the reference's own box, driven by sourcefile frames (fd23_box_configure_frames
# against the oracle's FileBox: PSV/psv_clean_double/box/box.cpp:10-210,
# material/material.cpp:39-99, main.cpp:153-200)
# ---------------------------------------------------------------------------
FBOX = dict(xs=120, zs=90, bw=32)


def _frames(nt, f64=True):
    """Frames of an explosive Ricker at the box centre in the homogeneous
    box medium (a run without a box), and that run's seismograms."""
    vp = np.full((NZ, NX), MED["vp"])
    vs = np.full((NZ, NX), MED["vs"])
    den = np.full((NZ, NX), MED["rho"])
    mat, _ = host.psv_material(vp, vs, den, H, DT, f64=True)
    packed = host.cpml_pack(_lib.WAVE_PSV, NX, NZ, NPML, DT, 1e-6, MED["vp"] / H, 12.0)
    sim = O.PSV(NX, NZ, NPML, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in mat])),
                unpack_cpml(packed, _lib.WAVE_PSV, NX, NZ, np.float64), dtype=np.float64)
    sim.dt_frames = DT
    stf = host.ricker(nt, DT, 12.0)
    M = host.psv_moment(0, 0, 0, 0, 0, H, DT, explosion=True).astype(np.float64)
    frames = point_source_frames(sim, nt, FBOX["xs"], FBOX["zs"], FBOX["bw"], stf, M)
    return frames, sim


def _file_models(hetero, f64):
    if hetero:
        vp, vs, den = layered_model(NX, NZ, 91, vmin=2000.0, vmax=4500.0)
    else:
        vp = np.full((NZ, NX), MED["vp"], np.float32)
        vs = np.full((NZ, NX), MED["vs"], np.float32)
        den = np.full((NZ, NX), MED["rho"], np.float32)
    if f64:
        vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
    mat, _ = host.psv_material(vp, vs, den, H, DT, f64=f64)
    packed = host.cpml_pack(_lib.WAVE_PSV, NX, NZ, NPML, DT, 1e-6, float(vp.max()) / H, 12.0)
    return mat, packed


def _file_oracle(mat, packed, frames, f64, nt, rx, rz):
    dt_ = np.float64 if f64 else np.float32
    sim = O.PSV(NX, NZ, NPML, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in mat])),
                unpack_cpml(packed, _lib.WAVE_PSV, NX, NZ, dt_), dtype=dt_)
    box = O.FileBox(sim, FBOX["xs"], FBOX["zs"], FBOX["bw"], frames, MED["vp"], MED["vs"], MED["rho"], H, DT)
    u, w = O.run_psv_box(sim, box, nt, rx, rz)
    return sim, box, u, w


def _file_gpu(mat, packed, frames, f64, nt, rx, rz, staged):
    storage = _lib.STORE_F64 if f64 else _lib.STORE_F32
    ctx = Context(_lib.WAVE_PSV, NX, NZ, NPML, storage=storage, max_frames=nt, max_rec=rx.size)
    ctx.set_material(mat)
    ctx.set_cpml(packed)
    for c in range(2):
        ctx.set_receivers(0, c, rx, rz)
    nforced = ctx.box_configure_frames(FBOX["xs"], FBOX["zs"], FBOX["bw"], frames, MED["vp"], MED["vs"],
                                       MED["rho"], H, DT)
    if staged:
        for it in range(nt):
            ctx.step_stress()
            ctx.step_velocity()
            ctx.record(it)
    else:
        ctx.run(0, nt)
    ctx.sync()
    return ctx, nforced


@pytest.mark.parametrize("f64,staged", [(True, False), (False, False), (True, True)])
def test_file_box_matches_oracle(f64, staged):
    """Layered model (most box cells forced to the source medium): the
    library's file-driven box against the oracle's, seismograms, residual
    grid and global fields."""
    print("case: file box f64=%s staged=%s" % (f64, staged), flush=True)
    nt = 300
    frames, _ = _frames(nt)
    rx = np.arange(2, NX - 2, 3, dtype=np.int32)
    rz = np.full(rx.size, 60, np.int32)  # 15 cells above the box
    mat, packed = _file_models(True, f64)
    sim, box, u, w = _file_oracle(mat, packed, frames, f64, nt, rx, rz)
    tol = 1e-10 if f64 else 1e-5
    ctx, nforced = _file_gpu(mat, packed, frames, f64, nt, rx, rz, staged)
    with ctx:
        res = ctx.box_residual()
        errs = {"seis U": rel(ctx.seismograms(0, 0, nt), u), "seis W": rel(ctx.seismograms(0, 1, nt), w)}
        vscale = max(np.linalg.norm(sim.U), np.linalg.norm(sim.W))
        sscale = max(np.linalg.norm(sim.Txx), np.linalg.norm(sim.Tzz), np.linalg.norm(sim.Txz))
        for name in ("U", "W", "Txx", "Tzz", "Txz"):
            sc = vscale if name in ("U", "W") else sscale
            errs["field " + name] = float(np.linalg.norm(ctx.field(name).astype(np.float64) - getattr(sim, name))
                                          / sc)
            errs["resid " + name] = float(np.linalg.norm(res[name].astype(np.float64) - box.r[name]) / sc)
    print("forced", nforced, box.nforced, errs, flush=True)
    assert nforced == box.nforced > 0
    assert np.abs(u).max() > 0
    bad = {k: v for k, v in errs.items() if not v < tol}
    assert not bad, bad


def test_file_box_is_transparent_on_the_gpu():
    """Homogeneous model, frames from the point-source run of the same
    scheme: the residual grid stays at rounding level and outside the box
    the fields equal the point-source run (FP64)."""
    print("case: file box transparent", flush=True)
    nt = 240
    frames, ref = _frames(nt)
    rx = np.array([10], np.int32)
    rz = np.array([10], np.int32)
    mat, packed = _file_models(False, True)
    ctx, nforced = _file_gpu(mat, packed, frames, True, nt, rx, rz, False)
    with ctx:
        res = ctx.box_residual()
        u = ctx.field("U")
    s0 = (FBOX["bw"] + 1) // 2 + 1 - FBOX["bw"]
    x0, z0, bw = FBOX["xs"] + s0, FBOX["zs"] + s0, FBOX["bw"]
    out = np.ones(u.shape, bool)
    out[z0:z0 + bw, x0:x0 + bw] = False
    peak = np.abs(frames[-1] - frames[-2]).max() / DT
    leak = max(np.abs(res[k]).max() for k in ("U", "W")) / peak
    e = rel(u[out], ref.U[out])
    print("forced", nforced, "leak", leak, "outside vs point source", e, flush=True)
    assert nforced == 0 and leak < 1e-9 and e < 1e-9


def test_file_box_frame_range_is_checked():
    nt = 40
    frames, _ = _frames(nt)
    mat, packed = _file_models(False, True)
    with Context(_lib.WAVE_PSV, NX, NZ, NPML, storage=_lib.STORE_F64, max_frames=nt + 10, max_rec=1) as ctx:
        ctx.set_material(mat)
        ctx.set_cpml(packed)
        ctx.box_configure_frames(FBOX["xs"], FBOX["zs"], FBOX["bw"], frames, MED["vp"], MED["vs"], MED["rho"], H,
                                 DT)
        with pytest.raises(_lib.Fd23Error) as e:
            ctx.run(0, nt + 1)  # needs nt+2 frames
        assert "frames" in str(e.value)
        ctx.run(0, nt)
        ctx.sync()
