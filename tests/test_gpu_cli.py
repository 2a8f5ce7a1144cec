"""The par-file driver end to end on the GPU, with the reference's own test
case PSV/clean_nobox/TEST/tt: t1 runs 500 steps and saves state000499, t2
restarts from it and saves state000999; snapshots every 100 steps.  Both
states and the U/W snapshots are compared with the reference's files
(1e-5 relative L2, as the library-level checkpoint test), and the ISIS
files with the raw seismograms they are written from."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fd23_amd", "bin", "nbpsv2d_topabsorb")
TT = os.path.join(ROOT, "tests", "golden", "ref", "psv_tt")


def _ref(name):
    return np.fromfile(os.path.join(TT, name), np.float32)


def _oracle_drift():
    """Relative L2 of the oracle's own chained 1000-step run (no restart
    from the reference's files) against the reference's late snapshots."""
    import oracle as O

    n, npml, h, dt = 96, 32, 0.3, 0.01
    vp, vs, den = O.mdl_raster(os.path.join(TT, "homo_6.8.mdl"), n, n, h)
    mat, _ = O.psv_material(vp, vs, den, h, dt)
    cp = O.psv_cpml(n, n, npml, 0.005, 1e-11, 30.0, 2.0)
    sim = O.PSV(n, n, npml, mat, cp, topabsorb=True)
    src, M = O.gaussian(-30), O.psv_moment(0, 0, 20, 90, 45, h, dt)
    snaps = {}
    O.run_psv(sim, 1000, src, M, 33, 33,
              snap_cb=lambda it, s: snaps.__setitem__(it, (s.U.copy(), s.W.copy())) if it in (500, 700, 900) else None)
    return {"U00500": rel(snaps[500][0].ravel(), _ref("U00500")), "U00700": rel(snaps[700][0].ravel(), _ref("U00700")),
            "W00900": rel(snaps[900][1].ravel(), _ref("W00900"))}


def test_tt_restart_pair_through_the_driver(tmp_path):
    print("case: driver t1/t2", flush=True)
    for f in ("t1", "t2", "station200"):
        shutil.copy(os.path.join(TT, f), tmp_path / f)
    shutil.copy(os.path.join(TT, "homo_6.8.mdl"), tmp_path / "homo.mdl")
    for par in ("t1", "t2"):
        r = subprocess.run([EXE, "par=" + par], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        print(r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0 and "finish" in r.stdout
    for name in ("state000499", "state000999", "U00100", "U00300", "W00300"):
        got = np.fromfile(tmp_path / name, np.float32)
        e = rel(got, _ref(name))
        print(name, e, flush=True)
        assert e < 1e-5, name
    # Late snapshots: once the wave has left through the C-PML the velocity
    # fields are small residuals, and FP32 runs that round differently from
    # the reference's drift apart relative to them (the oracle's own chained
    # run reaches 2.6e-5, 1.7e-4, 3.5e-4 at steps 500, 700, 900).  The driver
    # must stay within a few times the oracle's drift.
    drift = _oracle_drift()
    for name in ("U00500", "U00700", "W00900"):
        got = np.fromfile(tmp_path / name, np.float32)
        e = rel(got, _ref(name))
        print(name, e, "oracle", drift[name], flush=True)
        assert e < 3 * drift[name] + 1e-5, name
    # raw time-major seismograms (t2 overwrote t1's) and their ISIS copy;
    # the station sits at the corner (0,0) in the frozen C-PML edge, so as in
    # the reference's own output files of this case every sample is zero
    nrec = 2
    raw = np.fromfile(tmp_path / "0_U", np.float32).reshape(500, nrec)
    isis = np.fromfile(tmp_path / "0_U.isis", np.uint8).reshape(nrec, 84 + 4 * 500)
    for i in range(nrec):
        assert np.array_equal(isis[i, 84:].view(np.float32), raw[:, i])
    assert not raw.any()
    assert os.path.getsize(tmp_path / "0_W") == 500 * nrec * 4


def test_tt_with_first_arrival_window(tmp_path):
    """t1 for 200 steps with usepunch=1: a straight-ray travel-time file
    from the source at 6.8 km/s, before 0.1 s, after 0.3 s.  The saved
    state equals the oracle's run with the same window (1e-5)."""
    import oracle as O

    print("case: driver t1 with usepunch=1", flush=True)
    n, npml, h, dt, nt = 96, 32, 0.3, 0.01, 200
    z, x = np.mgrid[0:n, 0:n]
    times = (np.hypot(x - 33, z - 33) * h / 6.8).astype(np.float32)
    times.tofile(tmp_path / "tt.bin")
    par = open(os.path.join(TT, "t1")).read().replace("usepunch=0", "usepunch=1").replace("nt=500", f"nt={nt}")
    par += "\nbefore=0.1\nafter=0.3\ntimefile=tt.bin\n"
    (tmp_path / "tp").write_text(par)
    for f in ("station200",):
        shutil.copy(os.path.join(TT, f), tmp_path / f)
    shutil.copy(os.path.join(TT, "homo_6.8.mdl"), tmp_path / "homo.mdl")
    r = subprocess.run([EXE, "par=tp"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0 and "finish" in r.stdout
    vp, vs, den = O.mdl_raster(os.path.join(TT, "homo_6.8.mdl"), n, n, h)
    mat, _ = O.psv_material(vp, vs, den, h, dt)
    cp = O.psv_cpml(n, n, npml, 0.005, 1e-11, 30.0, 2.0)
    sim = O.PSV(n, n, npml, mat, cp, topabsorb=True)
    sim.set_window(times, 0.1, 0.3, dt)
    first, last = sim.window
    assert first.min() < 0 < first.max() and last.min() < nt  # blocks open and close in the run
    free = O.PSV(n, n, npml, mat, cp, topabsorb=True)
    src, M = O.gaussian(-30), O.psv_moment(0, 0, 20, 90, 45, h, dt)
    O.run_psv(sim, nt, src, M, 33, 33)
    O.run_psv(free, nt, src, M, 33, 33)
    got = np.fromfile(tmp_path / f"state{nt - 1:06d}", np.float32)
    want = np.frombuffer(sim.state_bytes(), np.float32)
    e = rel(got, want)
    print("state vs oracle with window", e, "vs without", rel(got, np.frombuffer(free.state_bytes(), np.float32)),
          flush=True)
    assert e < 1e-5
    assert rel(want, np.frombuffer(free.state_bytes(), np.float32)) > 1e-3


def test_fp64_tree_program_reproduces_the_reference_run(tmp_path):
    """psv2d_db on the FP64 tree's own test case (PSV/psv_clean_double/test:
    2048x1024, 2001 steps, the reference's 0_U, 0_W and 0_U.isis): the raw
    double seismograms within 1e-10 relative L2, the ISIS headers byte for
    byte (but the bytes the reference leaves uninitialised) and its samples
    equal to the raw file."""
    print("case: psv2d_db on psv_clean_double/test", flush=True)
    d = os.path.join(ROOT, "tests", "golden", "ref", "psv_fp64")
    for f in ("0.par", "homo.mdl", "station200"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    exe = os.path.join(ROOT, "fd23_amd", "bin", "psv2d_db")
    r = subprocess.run([exe, "par=0.par"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0 and "finish" in r.stdout
    for comp in ("U", "W"):
        got = np.fromfile(tmp_path / ("0_" + comp), np.float64)
        ref = np.fromfile(os.path.join(d, "0_" + comp), np.float64)
        e = rel(got, ref)
        print(comp, e, flush=True)
        assert got.size == ref.size == 2001 * 36 and e < 1e-10
    isis = np.frombuffer((tmp_path / "0_U.db.isis").read_bytes(), np.uint8).reshape(36, -1)
    ref = np.frombuffer(open(os.path.join(d, "0_U.isis"), "rb").read(), np.uint8).reshape(36, -1)
    assert isis.shape == ref.shape
    mask = np.ones(136, bool)
    mask[60:64] = mask[84:88] = mask[120:124] = False
    assert np.array_equal(isis[:, :136][:, mask], ref[:, :136][:, mask])
    raw = np.fromfile(tmp_path / "0_U", np.float64).reshape(2001, 36)
    for i in range(36):
        assert np.array_equal(isis[i, 136:].view(np.float64), raw[:, i])


def test_snapbox_windows_the_snapshots(tmp_path):
    """snapbox=x1,x2,z1,z2,it1,it2 (PSV/clean_nobox_new): the snapshot files
    in ./OUTPUT hold that window of the full snapshots, for steps in
    [it1, it2] only."""
    print("case: driver snapbox", flush=True)
    for f in ("t1", "station200"):
        shutil.copy(os.path.join(TT, f), tmp_path / f)
    shutil.copy(os.path.join(TT, "homo_6.8.mdl"), tmp_path / "homo.mdl")
    par = open(os.path.join(TT, "t1")).read().replace("nt=500", "nt=300")
    (tmp_path / "tb").write_text(par)
    r = subprocess.run([EXE, "par=tb"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-300:]
    full = {n: np.fromfile(tmp_path / n, np.float32).reshape(96, 96) for n in ("U00100", "W00200")}
    r = subprocess.run([EXE, "par=tb", "snapbox=10,50,20,60,100,200"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-300:]
    out = tmp_path / "OUTPUT"
    assert sorted(os.listdir(out)) == ["U00100", "U00200", "W00100", "W00200"]
    for n, f in full.items():
        win = np.fromfile(out / n, np.float32).reshape(41, 41)
        assert np.abs(f).max() > 0 and np.array_equal(win, f[20:61, 10:51]), n


def test_fp64_tree_program_with_the_file_box(tmp_path):
    """psv2d_db usebox=1 (the FP64 tree's box, PSV/psv_clean_double/
    main.cpp:153-200, keys param.cpp:79-85): a sourcefile of displacement
    frames written here from a point-source run of the oracle in the box
    medium; the program's raw seismograms against the oracle's FileBox run
    driven by the same file (1e-10), and against the point-source run
    itself, which the box reproduces outside it (1e-9)."""
    import oracle as O
    from fd23_amd import host
    from helpers import point_source_frames, unpack_cpml

    print("case: psv2d_db usebox=1", flush=True)
    nx, nz, npml, h, dt, nt = 256, 192, 20, 10.0, 8e-4, 300
    vp, vs, rho = 3000.0, 1732.0, 2300.0
    xs, zs, bw = 128, 100, 32
    nrec, ix0, iz0, idx = 40, 20, 60, 5
    mat, _ = host.psv_material(np.full((nz, nx), vp), np.full((nz, nx), vs), np.full((nz, nx), rho), h, dt,
                               f64=True)
    packed = host.cpml_pack(0, nx, nz, npml, dt, 1e-6, vp / h, 12.0)

    def sim():
        s = O.PSV(nx, nz, npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in mat])),
                  unpack_cpml(packed, 0, nx, nz, np.float64), dtype=np.float64)
        s.dt_frames = dt
        return s

    rx, rz = host.line_receivers(nrec, ix0, iz0, idx, 0)
    # frames from a point-source run; the same run's seismograms
    free = sim()
    stf = host.ricker(nt, dt, 12.0)
    M = host.psv_moment(0, 0, 0, 0, 0, h, dt, explosion=True).astype(np.float64)
    frames = point_source_frames(free, nt, xs, zs, bw, stf, M)
    free2 = sim()
    fu, fw = O.run_psv(free2, nt, stf, M, xs, zs, rx=rx, rz=rz)
    frames.tofile(tmp_path / "box.bin")
    # the oracle's file box on the same inputs
    osim = sim()
    box = O.FileBox(osim, xs, zs, bw, frames, vp, vs, rho, h, dt)
    ou, ow = O.run_psv_box(osim, box, nt, rx, rz)
    (tmp_path / "box.mdl").write_text("BACK %g %g %g\n" % (vp, vs, rho))
    par = "\n".join([
        "usebox=1", "usepunch=0", "ngpu=1", "nx=%d" % nx, "nz=%d" % nz, "nt=%d" % nt, "h=%g" % h, "dt=%g" % dt,
        "srcden=%g" % rho, "srcsvel=%g" % vs, "srcpvel=%g" % vp, "sourcefile=box.bin", "bw=%d" % bw,
        "sourcetime=Gaussian", "alpha=-5", "strike=0", "dip=0", "rake=0", "azimuth=0",
        "xs=%d" % xs, "zs=%d" % zs, "npml=%d" % npml, "pml_dt=%g" % dt, "pml_v=%g" % (vp / h), "pml_r=1e-6",
        "pml_fc=12", "itrecord=1", "output=bx", "nrec=%d" % nrec, "ixrec0=%d" % ix0, "izrec0_u=%d" % iz0,
        "izrec0_w=%d" % iz0, "idxrec=%d" % idx, "idzrec=0", "model=box.mdl", "ntsnap=100000", ""])
    (tmp_path / "box.par").write_text(par)
    exe = os.path.join(ROOT, "fd23_amd", "bin", "psv2d_db")
    r = subprocess.run([exe, "par=box.par"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0 and "finish" in r.stdout
    for comp, o, f in (("U", ou, fu), ("W", ow, fw)):
        got = np.fromfile(tmp_path / ("bx_" + comp), np.float64).reshape(nt, -1)
        e_box, e_free = rel(got, o), rel(got, f)
        print(comp, "vs oracle file box", e_box, "vs point-source run", e_free, flush=True)
        assert np.abs(o).max() > 0 and e_box < 1e-10 and e_free < 1e-9
