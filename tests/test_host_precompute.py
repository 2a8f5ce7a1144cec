"""Host precompute (library C++ and the oracle) against vectors written by the
reference's own host code (tests/golden/make_ref_vectors.py runs
PSV/clean_nobox and SH material/cpml/waveform/model/getpar sources compiled
in oracle/_ref).  CPU only: these library functions make no GPU calls."""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from fd23_amd import _lib, host
from helpers import rel, unpack_cpml

D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_host")
PSV_CASES = ["psv_layered", "psv_water", "psv_trap"]


def load(name):
    return dict(np.load(os.path.join(D, name + ".npz")))


def _write_case(d, g):
    for k in ("vp", "vs", "den"):
        g[k].tofile(os.path.join(d, "mod." + k))
    open(os.path.join(d, "case.par"), "w").write(g["par"].tobytes().decode())
    open(os.path.join(d, "nested.par"), "w").write(g["nested"].tobytes().decode())


@pytest.mark.parametrize("case", PSV_CASES)
def test_par_file_semantics(case):
    """Comments, several entries per line, quotes, $(name), nested par=,
    backslash continuation, last occurrence wins, argv after par=."""
    g = load(case)
    with tempfile.TemporaryDirectory() as d:
        _write_case(d, g)
        cwd = os.getcwd()
        os.chdir(d)
        try:
            p = host.Par(["prog", "par=case.par", "pml_fc=2.5"])
            ints = [p.int(k) for k in ("nx", "nz", "npml", "xs", "zs", "nt", "itrecord")]
            flts = [p.float(k) for k in ("h", "dt", "strike", "dip", "rake", "azimuth", "pml_r", "pml_v")]
            assert p.float("pml_fc") == np.float32(2.5)
            assert p.str("model") == "mod" and p.int("missing_key") is None
        finally:
            os.chdir(cwd)
    assert ints == list(g["ints"][:7])
    assert np.array_equal(np.array(flts, np.float32), g["floats"])


@pytest.mark.parametrize("case", PSV_CASES)
def test_psv_material_lib_and_oracle(case):
    g = load(case)
    nx, nz = int(g["ints"][0]), int(g["ints"][1])
    h, dt = float(g["floats"][0]), float(g["floats"][1])
    mat, stab = host.psv_material(g["vp"], g["vs"], g["den"], h, dt)
    omat, _ = O.psv_material(g["vp"], g["vs"], g["den"], h, dt)
    for a, k in zip(mat, ("BU", "BW", "MU", "MUA", "LAM")):
        assert a.shape == (nz, nx)
        np.testing.assert_allclose(a, g[k], rtol=2e-7, atol=0, err_msg="lib " + k)
        np.testing.assert_allclose(omat[k], g[k], rtol=2e-7, atol=0, err_msg="oracle " + k)


def test_mua_quirk_is_exercised():
    """The layered fixture has interfaces where MUA's repeated m2m3m4 term
    differs from the true harmonic mean, so the check above pins it."""
    g = load("psv_layered")
    h, dt = float(g["floats"][0]), float(g["floats"][1])
    mu = (g["vs"].astype(np.float64) ** 2 * g["den"])
    m1, m2, m3, m4 = mu[1:-1, :-1], mu[1:-1, 1:], mu[:-2, :-1], mu[:-2, 1:]
    true = 4 * m1 * m2 * m3 * m4 / (m2 * m3 * m4 + m1 * m3 * m4 + m1 * m2 * m4 + m1 * m2 * m3) * dt / h
    ref = g["MUA"][1:-1, :-1].astype(np.float64)
    assert np.abs(true - ref).max() / np.abs(ref).max() > 1e-4


@pytest.mark.parametrize("case", PSV_CASES)
def test_psv_cpml_profiles(case):
    g = load(case)
    nx, nz, npml = (int(v) for v in g["ints"][:3])
    fl = g["floats"]
    pml_dt, pml_r, pml_v, pml_fc = 0.01, float(fl[6]), float(fl[7]), 2.5
    packed = host.cpml_pack(_lib.WAVE_PSV, nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc)
    cp = unpack_cpml(packed, _lib.WAVE_PSV, nx, nz, np.float64)
    ocp = O.psv_cpml(nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc)
    for name in cp:
        for i, a in enumerate("bck"):
            ref = g["%s_%s" % (a, name)]
            # library computes in double (rounded once); the reference in float
            np.testing.assert_allclose(cp[name][i], ref, rtol=3e-6, atol=1e-12, err_msg="lib %s %s" % (a, name))
            np.testing.assert_allclose(ocp[name][i], ref, rtol=1e-6, atol=1e-12, err_msg="oracle %s %s" % (a, name))


def test_cpml_profile_properties():
    b, c, k = host.cpml_profile(100, 20, 0.5, 1e-3, 1e-6, 800.0, 10.0)
    assert np.all(b[:4] == 0) and np.all(b[-5:] == 0) and np.all(k[:4] == 0) and np.all(k[-5:] == 0)
    inner = slice(25, 75)
    assert np.allclose(k[inner], 1.0) and np.allclose(c[inner], 0.0)
    assert np.all((b[4:-5] > 0) & (b[4:-5] <= 1)) and np.all(c[4:-5] <= 0)
    # damping grows toward both edges
    assert b[5] < b[15] and b[-6] < b[-16]


@pytest.mark.parametrize("case", PSV_CASES)
def test_source_time_functions(case):
    g = load(case)
    txt = g["par"].tobytes().decode()
    if "srctime=T" in txt:
        lib = host.trapezoid(0.01, 0.05, 0.03, 0.08)
        orc = O.trapezoid(0.01, 0.05, 0.03, 0.08)
    else:
        alpha = float(txt.split("alpha=")[1].split()[0])
        lib = host.gaussian(alpha)
        orc = O.gaussian(alpha)
    np.testing.assert_allclose(lib, g["src"], rtol=1e-7, atol=0)
    np.testing.assert_allclose(orc, g["src"], rtol=1e-7, atol=0)


def test_gaussian_alpha_truncates_to_int():
    assert np.array_equal(host.gaussian(-7.9), host.gaussian(-7))
    s = host.gaussian(-5)
    assert len(s) == 31 and abs(s.sum() - 1) < 1e-6


def test_mdl_model_and_material():
    g = load("psv_mdl")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.mdl")
        open(p, "w").write(g["mdl"].tobytes().decode())
        vp, vs, den = host.model_read(p, 64, 48, 0.25)
        ovp, ovs, oden = O.mdl_raster(p, 64, 48, 0.25)
    assert np.array_equal(vp, ovp) and np.array_equal(vs, ovs) and np.array_equal(den, oden)
    mat, _ = host.psv_material(vp, vs, den, 0.25, 0.01)
    for a, k in zip(mat, ("BU", "BW", "MU", "MUA", "LAM")):
        np.testing.assert_allclose(a, g[k], rtol=2e-7, atol=0, err_msg=k)


def test_sh_material_and_cpml():
    g = load("sh_layered")
    mat, _ = host.sh_material(g["vs"], g["den"], 0.2, 0.01)
    omat, _ = O.sh_material(g["vs"], g["den"], 0.2, 0.01)
    for a, k in zip(mat, ("BV", "MUX", "MUZ")):
        np.testing.assert_allclose(a, g[k], rtol=2e-7, atol=0, err_msg="lib " + k)
        np.testing.assert_allclose(omat[k], g[k], rtol=2e-7, atol=0, err_msg="oracle " + k)
    nz, nx = g["vs"].shape
    packed = host.cpml_pack(_lib.WAVE_SH, nx, nz, 7, 0.01, 1e-6, 30.0, 2.0)
    cp = unpack_cpml(packed, _lib.WAVE_SH, nx, nz, np.float64)
    ocp = O.sh_cpml(nx, nz, 7, 0.01, 1e-6, 30.0, 2.0)
    for name in cp:
        for i, a in enumerate("bck"):
            ref = g["%s_%s" % (a, name)]
            np.testing.assert_allclose(cp[name][i], ref, rtol=3e-6, atol=1e-12)
            np.testing.assert_allclose(ocp[name][i], ref, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(host.gaussian(-6), g["src"], rtol=1e-7)


def test_stability_failure_is_an_error():
    vp = np.full((20, 20), 6.0, np.float32)
    with pytest.raises(_lib.Fd23Error) as e:
        host.psv_material(vp, vp / 1.7, vp * 0 + 3, 0.01, 0.01)
    assert e.value.code == _lib.ESTAB


def test_missing_model_and_par_file_are_errors():
    with pytest.raises(_lib.Fd23Error) as e:
        host.model_read("/nonexistent/model", 10, 10, 1.0)
    assert e.value.code == _lib.EIO
    with pytest.raises(_lib.Fd23Error):
        host.Par(["prog", "par=/nonexistent.par"])


def test_isis_layout_matches_reference_file():
    """Data section of the FP32 ISIS writer = raw file transposed, 84-byte
    headers with the reference's fields (PSV/clean_nobox/record/isis.cpp)."""
    rng = np.random.default_rng(0)
    raw = rng.standard_normal((7, 4)).astype(np.float32)
    rx = np.array([11, 10, 13, 12], np.int32)
    rz = np.array([0, 0, 2, 2], np.int32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "out_U.isis")
        host.isis_write(p, raw, rx, rz, 0.5, 0.0, 0.3, 0.01, 1, 512, 4)
        b = np.fromfile(p, np.uint8)
    assert b.size == 4 * (84 + 4 * 7)
    for i in range(4):
        rec = b[i * (84 + 28):(i + 1) * (84 + 28)]
        hdr = rec[:84]
        f = hdr.view(np.float32)
        iv = hdr.view(np.int32)
        assert np.isclose(f[3], (rx[i] + 0.5) * np.float32(0.3)) and iv[7] == 7
        assert iv[9] == 512 and iv[10] == rx[i] and iv[-2] == 123456789 and iv[-3] == 123456789
        assert np.array_equal(rec[84:].view(np.float32), raw[:, i])


def test_station_file_and_line_receivers():
    """The library's receiver lists equal the oracle's (whose station parsing
    reproduces the reference's FP64 seismogram fixture), pairs at x, x-1."""
    import oracle as O

    path = os.path.join(os.path.dirname(D), "ref", "psv_fp64", "station200")
    rx, rz = host.station_read(path)
    ox, oz = O.station_receivers(path)
    assert np.array_equal(rx, ox) and np.array_equal(rz, oz)
    assert rx.size == 36 and (rz[0], rx[0], rx[1]) == (711, 529, 528)
    lx, lz = host.line_receivers(5, 513, 4, 2, 1)
    ox, oz = O.line_receivers(5, 513, 4, 2, 1)
    assert np.array_equal(lx, ox) and np.array_equal(lz, oz)
    with tempfile.TemporaryDirectory() as d:
        bad = os.path.join(d, "odd")
        open(bad, "w").write("1 2 3\n")
        with pytest.raises(_lib.Fd23Error):
            host.station_read(bad)
    with pytest.raises(_lib.Fd23Error):
        host.station_read("/nonexistent/stations")


# ---------------------------------------------------------------------------
# the FP64 tree's host steps (PSV/psv_clean_double): model, wavelets, moment,
# and its ISIS file against the reference's own output of that tree
# ---------------------------------------------------------------------------
FP64_DIR = os.path.join(os.path.dirname(D), "ref", "psv_fp64")


def test_fp64_model_wavelet_moment_match_oracle():
    for path, nx, nz, h in ((os.path.join(FP64_DIR, "homo.mdl"), 2048, 1024, 0.3),):
        got = host.model_read_f64(path, nx, nz, h)
        want = O.mdl_raster(path, nx, nz, h, np.float64)
        for a, b in zip(got, want):
            assert a.dtype == np.float64 and np.array_equal(a, b)
        # FP64 values are the decimal strings parsed as double, not widened floats
        assert got[1].max() == 4.0 and 6.8 in got[0] and np.float64(np.float32(6.8)) not in got[0]
    for a in (-5, -30, 3):
        assert np.array_equal(host.gaussian_f64(a), O.gaussian(a, np.float64)), a
    assert abs(host.gaussian_f64(-5).sum() - 1.0) < 1e-15
    for kind in (0, 1):
        M = host.psv_moment_f64(kind, 0.0, 20.0, 90.0, 45.0, 0.3, 0.01)
        assert np.array_equal(M, O.psv_moment(kind, 0.0, 20.0, 90.0, 45.0, 0.3, 0.01, np.float64)), kind
    # trapezoid: unit area in samples (int s dt = 1), the rise, flat and fall lengths
    s = host.trapezoid_f64(0.01, 0.3, 0.0, 0.3)
    assert len(s) == 60 and abs(s.sum() - 1.0) < 0.05 and s[0] == 0.0 and s.argmax() in (29, 30)


def test_fp64_isis_reproduces_the_reference_file(tmp_path):
    """PSV/psv_clean_double/test/0_U.isis: 36 traces of 2001 doubles behind
    136-byte headers.  Written again from the raw file (0_U) with the run's
    parameters (0.par), every byte agrees except the ones the reference
    leaves uninitialised (the padding after nt and after status, user[0])."""
    raw = np.fromfile(os.path.join(FP64_DIR, "0_U"), np.float64).reshape(2001, 36)
    rx, rz = O.station_receivers(os.path.join(FP64_DIR, "station200"))
    out = tmp_path / "0_U.db.isis"
    host.isis_write_f64(str(out), raw, rx, rz, 0.5, 0.0, 0.3, 0.01, 1, 1024, 100, 0.25, -0.25, 0)
    got = np.frombuffer(out.read_bytes(), np.uint8).reshape(36, -1)
    ref = np.frombuffer(open(os.path.join(FP64_DIR, "0_U.isis"), "rb").read(), np.uint8).reshape(36, -1)
    assert got.shape == ref.shape == (36, 136 + 8 * 2001)
    mask = np.ones(got.shape[1], bool)
    mask[60:64] = mask[84:88] = mask[120:124] = False
    assert np.array_equal(got[:, mask], ref[:, mask])
