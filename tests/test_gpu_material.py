"""Material computed on the device (fd23_set_model) against the host
precompute (fd23_psv_material[_f64], fd23_sh_material, themselves pinned to
the reference's own host code by tests/golden/ref_host): every array
bitwise, the same stability number, the same FP16 scale estimate, and
runs from either material equal bitwise.

The models have water cells (vs = 0) inside the 5-cell edge band, where the
reference replaces mu by vp^2 rho / 3, and outside it, where it does not;
sizes include a non-multiple of the 256-thread block and a two-row grid
edge case for the wrapping neighbours.
"""
import numpy as np
import pytest

from fd23_amd import _lib, host
from fd23_amd.core import Context
from helpers import Problem, layered_model

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _model(nx, nz, seed, f64=False):
    vp, vs, den = layered_model(nx, nz, seed)
    rng = np.random.default_rng(seed)
    # water: a column strip at the left edge, bottom rows, and interior cells
    vs[:, :3] = 0.0
    vs[-2:, nx // 3:nx // 2] = 0.0
    idx = rng.integers(0, nx * nz, 40)
    vs.ravel()[idx] = 0.0
    if f64:
        vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
    return vp, vs, den


@pytest.mark.parametrize("wave,storage,nx,nz", [
    (_lib.WAVE_PSV, _lib.STORE_F32, 150, 90),
    (_lib.WAVE_PSV, _lib.STORE_F32, 333, 17),
    (_lib.WAVE_PSV, _lib.STORE_F64, 150, 90),
    (_lib.WAVE_PSV, _lib.STORE_F16, 128, 64),
    (_lib.WAVE_SH, _lib.STORE_F32, 131, 77),
    (_lib.WAVE_SH, _lib.STORE_BF16, 64, 16),
])
def test_device_material_equals_host(wave, storage, nx, nz):
    _print("case: material", wave, storage, nx, nz)
    h, dt = 10.0, 7e-4
    f64 = storage == _lib.STORE_F64
    vp, vs, den = _model(nx, nz, nx + nz, f64)
    if wave == _lib.WAVE_PSV:
        mat, stab = host.psv_material(vp, vs, den, h, dt, f64=f64)
    else:
        mat, stab = host.sh_material(vs, den, h, dt)
    with Context(wave, nx, nz, 8, storage=storage) as ctx:
        s = ctx.set_model(vp, vs, den, h, dt)
        got = [ctx.material(i) for i in range(len(mat))]
    _print("stab", s, stab)
    assert s == stab
    for i, (g, m) in enumerate(zip(got, mat)):
        assert g.dtype == m.dtype
        same = g.view(np.uint64 if f64 else np.uint32) == m.view(np.uint64 if f64 else np.uint32)
        assert same.all(), (i, np.argwhere(~same)[:5])


@pytest.mark.parametrize("wave", [_lib.WAVE_PSV, _lib.WAVE_SH])
def test_runs_from_device_material_equal_host(wave):
    """Seismograms, fields and the FP16 scale estimate (from the largest
    buoyancy, reduced on the device) equal those of the host material."""
    _print("case: run from device material", wave)
    for storage in (_lib.STORE_F32, _lib.STORE_F16):
        p = Problem(wave, 128, 64, 10, 10.0, 7e-4, 300, seed=21, sources=[(50, 8)], rec_z=3)
        vp, vs, den = layered_model(p.nx, p.nz, 21, vmax=6000.0)
        out = []
        for dev in (False, True):
            with p.context(storage=storage) as ctx:
                if dev:
                    ctx.set_model(vp, vs, den, p.h, p.dt)
                ctx.run(0, p.nt, 1)
                sc = ctx.get_scales(13 if wave == _lib.WAVE_PSV else 7)[0] if storage != _lib.STORE_F32 else None
                out.append((ctx.seismograms(0, 0, p.nt), ctx.field("U" if wave == _lib.WAVE_PSV else "V"), sc))
        (a, fa, sa), (b, fb, sb) = out
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b) and np.array_equal(fa, fb)
        if sa is not None:
            assert np.array_equal(sa, sb)


def test_unstable_model_reports_and_sets():
    _print("case: material stability")
    vp, vs, den = _model(96, 48, 3)
    with Context(_lib.WAVE_PSV, 96, 48, 8) as ctx:
        with pytest.raises(_lib.Fd23Error, match="stability"):
            ctx.set_model(vp, vs, den, 10.0, 5e-3)
    with Context(_lib.WAVE_SH, 96, 48, 8, storage=_lib.STORE_F64) as ctx:
        with pytest.raises(_lib.Fd23Error, match="SH material"):
            ctx.set_model(None, vs.astype(np.float64), den.astype(np.float64), 10.0, 7e-4)


SYNTH_MDL = """BACK 3.0 1.7 2.2
DEF A 1.0 1.0
DEF B 9.0 1.0
# a triangle with a horizontal base on grid rows, vertices on grid points
POLY 4.0 -0.5 2.4
\tA B
\t5.0 6.0
# overlapping later polygon (wins), water (vs 0)
POLY 1.5 0.0 1.0
\t4.0 0.0
\t12.0 0.0
\t12.0 3.0
\t4.0 3.0
# a sliver with a vertical edge through cell centres and one off the grid
POLY 5.5 3.1 -0.45
\t2.5 4.0
\t2.5 9.5
\t7.25 7.75
"""


@pytest.mark.parametrize("case", ["fixture", "tt", "synthetic", "synthetic_f64", "synthetic_sh", "raster"])
def test_model_file_on_device_equals_host(tmp_path, case):
    """fd23_set_model_file: the .mdl polygons rasterised on the device, then
    the material, bitwise equal to fd23_model_read + the host precompute;
    raster files go through the host reader."""
    import os

    _print("case: model file", case)
    wave, storage, h, dt = _lib.WAVE_PSV, _lib.STORE_F32, 0.25, 0.01
    if case == "fixture":
        g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_host", "psv_mdl.npz"))
        path, nx, nz = tmp_path / "m.mdl", 64, 48
        path.write_text(g["mdl"].tobytes().decode())
    elif case == "tt":
        path, nx, nz, h = os.path.join(os.path.dirname(__file__), "golden", "ref", "psv_tt", "homo_6.8.mdl"), 96, 96, 0.3
    elif case == "raster":
        nx, nz = 70, 40
        vp, vs, den = layered_model(nx, nz, 9)
        for ext, a in (("vp", vp), ("vs", vs), ("den", den)):
            a.astype(np.float32).tofile(tmp_path / ("r." + ext))
        path, h, dt = tmp_path / "r", 10.0, 7e-4
    else:
        path, nx, nz = tmp_path / "s.mdl", 56, 44
        path.write_text(SYNTH_MDL)
        h, dt = 0.25, 0.005
        if case == "synthetic_f64":
            storage = _lib.STORE_F64
        if case == "synthetic_sh":
            wave = _lib.WAVE_SH
    vp, vs, den = host.model_read(str(path), nx, nz, h)
    f64 = storage == _lib.STORE_F64
    if wave == _lib.WAVE_PSV:
        mat, stab = host.psv_material(*((a.astype(np.float64) for a in (vp, vs, den)) if f64 else (vp, vs, den)),
                                      h, dt, f64=f64)
    else:
        mat, stab = host.sh_material(vs, den, h, dt)
    if case.startswith("synthetic"):
        # every branch of the polygon test is taken somewhere on this grid
        assert len(np.unique(vp)) == 4
    with Context(wave, nx, nz, 8, storage=storage) as ctx:
        s = ctx.set_model_file(path, h, dt)
        got = [ctx.material(i) for i in range(len(mat))]
    assert s == stab
    for i, (a, b) in enumerate(zip(got, mat)):
        bits = np.uint64 if f64 else np.uint32
        assert np.array_equal(a.view(bits), b.view(bits)), (case, i)
