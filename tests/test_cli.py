"""The par-file driver (fd23_amd/bin, the reference's nbpsv2d / sh2d
interface) on the CPU: it reads the reference's own par files, model and
station files, and stops with an error before any device work when a key
is missing or an input file cannot be read.  With no GPU the run ends at
context creation, after every host step (par, model, material, C-PML,
wavelet, moment tensor, receivers) has succeeded."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fd23_amd", "bin")
TT = os.path.join(ROOT, "tests", "golden", "ref", "psv_tt")


def _run(prog, args, cwd):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    return subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


@pytest.fixture()
def ttdir(tmp_path):
    shutil.copy(os.path.join(TT, "t1"), tmp_path / "t1")
    shutil.copy(os.path.join(TT, "homo_6.8.mdl"), tmp_path / "homo.mdl")
    shutil.copy(os.path.join(TT, "station200"), tmp_path / "station200")
    return tmp_path


def test_reference_par_file_reaches_the_device(ttdir):
    r = _run("nbpsv2d_topabsorb", ["par=t1"], ttdir)
    try:
        import torch

        gpu = torch.cuda.is_available()
    except Exception:
        gpu = False
    if gpu:
        pytest.skip("a GPU is present: the GPU test runs this case to the end")
    assert r.returncode != 0
    assert "fd23_ctx_create" in r.stderr, r.stderr


def test_missing_key_is_reported(ttdir):
    # getpar exits with GETPAR_ERROR = 100 on a missing mstpar key and on a
    # par file it cannot open (PSV/clean_nobox/getpar/getpar.c:82,442,646)
    r = _run("sh2d", ["par=t1"], ttdir)
    assert r.returncode == 100 and "must specify value for 'nrec'" in r.stderr
    r = _run("nbpsv2d", ["par=no_such_file"], ttdir)
    assert r.returncode == 100 and "no_such_file" in r.stderr


def test_punch_keys(ttdir):
    # usepunch=1 makes before, after and timefile mandatory (param.cpp:187-191)
    r = _run("nbpsv2d", ["par=t1", "usepunch=1"], ttdir)
    assert r.returncode != 0 and "must specify value for 'before'" in r.stderr
    r = _run("nbpsv2d", ["par=t1", "usepunch=1", "before=0.1", "after=0.2", "timefile=none.bin"], ttdir)
    assert r.returncode != 0 and "cannot read" in r.stderr and "none.bin" in r.stderr
    # a travel-time file one value short
    (ttdir / "short.bin").write_bytes(b"\0" * (4 * (96 * 96 - 1)))
    r = _run("nbpsv2d", ["par=t1", "usepunch=1", "before=0.1", "after=0.2", "timefile=short.bin"], ttdir)
    assert r.returncode != 0 and "cannot read" in r.stderr


def test_missing_model_is_an_error(ttdir):
    r = _run("nbpsv2d", ["par=t1", "model=none.mdl"], ttdir)
    assert r.returncode != 0 and "fd23_model_read" in r.stderr


@pytest.fixture()
def fp64dir(tmp_path):
    d = os.path.join(ROOT, "tests", "golden", "ref", "psv_fp64")
    for f in ("0.par", "homo.mdl", "station200"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    return tmp_path


def test_fp64_tree_par_file(fp64dir):
    """psv2d_db reads the FP64 tree's par file (PSV/psv_clean_double/test/
    0.par, its keys: usebox, sourcetime, stype, src_shift_*) through every
    host step; without a GPU it stops at context creation."""
    r = _run("psv2d_db", ["par=0.par"], fp64dir)
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present: the GPU test runs this case to the end")
    except Exception:
        pass
    assert r.returncode != 0 and "fd23_ctx_create" in r.stderr, r.stderr
    # usebox=1 makes the box keys mandatory (param.cpp:79-85)
    r = _run("psv2d_db", ["par=0.par", "usebox=1"], fp64dir)
    assert r.returncode == 100 and "must specify value for 'srcden'" in r.stderr
    box = ["usebox=1", "srcden=3.0", "srcsvel=4.0", "srcpvel=6.8", "bw=24"]
    r = _run("psv2d_db", ["par=0.par", "sourcefile=none.bin"] + box + ["zs=14"], fp64dir)
    assert r.returncode != 0 and "zs 14 is too near to the boundary" in r.stderr
    # a source file shorter than nt+1 frames (here 3 frames for nt=2001)
    (fp64dir / "short.bin").write_bytes(b"\0" * (8 * 2 * 32 * 32 * 3))
    r = _run("psv2d_db", ["par=0.par", "sourcefile=short.bin"] + box, fp64dir)
    assert r.returncode != 0 and "cannot read 2002 box source frames" in r.stderr
    (fp64dir / "nobox.par").write_text(open(fp64dir / "0.par").read().replace("usebox=0", ""))
    r = _run("psv2d_db", ["par=nobox.par"], fp64dir)
    assert r.returncode != 0 and "must specify value for 'usebox'" in r.stderr


def test_snapbox_is_checked(ttdir):
    r = _run("nbpsv2d", ["par=t1", "snapbox=1,2,3"], ttdir)
    assert r.returncode != 0 and "snapbox should be six integer" in r.stderr
    r = _run("nbpsv2d", ["par=t1", "snapbox=0,96,0,10,0,100"], ttdir)
    assert r.returncode != 0 and "outside" in r.stderr


def test_sh_output_material(tmp_path):
    """sh2d output_material=1 writes the .mdl model's vs and den
    ({model}_VS, _DEN: two int sizes, then nx*nz floats) before any device
    work; they equal the library's model reader."""
    import numpy as np

    from fd23_amd import host

    shutil.copy(os.path.join(TT, "homo_6.8.mdl"), tmp_path / "homo.mdl")
    par = ("nx=96 nz=96 nt=10 h=0.3 dt=0.01 usepunch=0 ngpu=1 sourcetime=Gaussian alpha=-30 strike=0 dip=20 "
           "rake=90 azimuth=45 xs=33 zs=33 npml=32 pml_dt=0.005 pml_v=30 pml_r=1e-11 pml_fc=2 itrecord=1 "
           "output=o nrec=2 ixrec0=10 izrec0_v=10 idxrec=1 idzrec=0 model=homo.mdl ntsnap=100 "
           "output_material=1").split()
    _run("sh2d", par, tmp_path)
    _, vs, den = host.model_read(str(tmp_path / "homo.mdl"), 96, 96, 0.3)
    for name, ref in (("homo.mdl_VS", vs), ("homo.mdl_DEN", den)):
        b = (tmp_path / name).read_bytes()
        assert np.frombuffer(b[:8], np.int32).tolist() == [96, 96]
        assert np.array_equal(np.frombuffer(b[8:], np.float32).reshape(96, 96), ref)
