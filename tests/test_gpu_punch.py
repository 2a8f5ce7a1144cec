"""First-arrival window ("punch", SH/slide/cu_slide.cu:7-61 and the
kernels' CHECK_IF_NEED_UPDATE, SH/fd/cu_fd.cu:18-25) on the GPU against the
oracle with the same window.

The window comes from straight-ray travel times at the model's fastest P
speed with before = 10 ms and after = 60 ms, so over the 400 steps blocks
open as the wave front approaches and close behind it: cells are frozen
for part of the run, tiles with no open block go idle, and the top block
row is open in some columns and closed in others (the SH surface
imaging).  Tolerances as the other parity tests: 1e-5 relative L2 for
FP32, 1e-10 for FP64; every tile shape gives the same result bitwise.
"""
import numpy as np
import pytest

from fd23_amd import _lib
from helpers import Problem, rel

pytestmark = pytest.mark.gpu
BEFORE, AFTER = 0.01, 0.06


def _print(*a):
    print(*a, flush=True)


def _run(p, window, tile=0, storage=None):
    with p.context(storage=storage) as ctx:
        ctx.tile = tile
        if window is not None:
            ctx.set_punch(*window, p.dt)
        ctx.run(0, p.nt, 1)
        ncomp = 2 if p.wave == _lib.WAVE_PSV else 1
        seis = [[ctx.seismograms(s, c, p.nt) for c in range(ncomp)] for s in range(p.nshots)]
        names = ("U", "W", "Txx", "Txz", "Tzz") if p.wave == _lib.WAVE_PSV else ("V", "Tx", "Tz")
        fields = [{n: ctx.field(n, s) for n in names} for s in range(p.nshots)]
    return seis, fields


def _check(p, seis, fields, window, tol):
    for s in range(p.nshots):
        sim, ref = p.oracle(s, window=window)
        for c, r in enumerate(ref):
            e = rel(seis[s][c], r)
            _print("shot", s, "comp", c, "seis", e)
            assert e < tol
        for n, f in fields[s].items():
            e = rel(f, getattr(sim, n))
            _print("shot", s, n, e)
            assert e < tol


def _window(p):
    return p.travel_times(0), BEFORE, AFTER


def _psv_problem(f64=False):
    # two shots (one window, from shot 0's source) so the multi-shot tiles
    # cover a shot group
    return Problem(_lib.WAVE_PSV, 128, 64, 10, 10.0, 7e-4, 400, seed=11, sources=[(40, 20), (90, 30)],
                   rec_z=3, f64=f64)


PSV_TILES = [0, 1, 2, 3, 14, 15]  # RZ 2/4/8, SPB 2, split RZ 2/4


@pytest.mark.parametrize("tile", PSV_TILES)
def test_psv_window_vs_oracle(tile):
    _print("case: psv punch tile", tile)
    p = _psv_problem()
    w = _window(p)
    first_open = (w[0].min() - BEFORE) / p.dt
    assert first_open < 0  # the source block is open from the start
    seis, fields = _run(p, w, tile)
    _check(p, seis, fields, w, 1e-5)
    # the window changes the result
    free, _ = _run(p, None, tile)
    d = rel(seis[0][0], free[0][0])
    _print("window vs none", d)
    assert d > 1e-3


def test_psv_window_top_absorbing_p_source():
    _print("case: psv punch top C-PML, p source")
    p = Problem(_lib.WAVE_PSV, 96, 48, 8, 10.0, 7e-4, 400, seed=12, topabsorb=True, kind=_lib.SRC_P,
                sources=[(30, 20)], rec_z=11)
    w = _window(p)
    seis, fields = _run(p, w)
    _check(p, seis, fields, w, 1e-5)


def test_psv_window_fp64():
    _print("case: psv punch fp64")
    p = _psv_problem(f64=True)
    w = _window(p)
    seis, fields = _run(p, w)
    _check(p, seis, fields, w, 1e-10)


@pytest.mark.parametrize("tile", [0, 1])
def test_sh_window_vs_oracle(tile):
    _print("case: sh punch tile", tile)
    p = Problem(_lib.WAVE_SH, 128, 64, 10, 10.0, 7e-4, 400, seed=13, kind=_lib.SRC_FORCE,
                sources=[(40, 12)], rec_z=2)
    w = _window(p)
    seis, fields = _run(p, w, tile)
    _check(p, seis, fields, w, 1e-5)
    free, _ = _run(p, None, tile)
    assert rel(seis[0][0], free[0][0]) > 1e-3


def test_sh_window_top_absorbing():
    _print("case: sh punch top C-PML")
    p = Problem(_lib.WAVE_SH, 96, 48, 8, 10.0, 7e-4, 400, seed=14, topabsorb=True, kind=_lib.SRC_LINE,
                sources=[(60, 20)], rec_z=10)
    w = _window(p)
    seis, fields = _run(p, w)
    _check(p, seis, fields, w, 1e-5)


@pytest.mark.parametrize("wave", [_lib.WAVE_PSV, _lib.WAVE_SH])
@pytest.mark.parametrize("tile", [0, 14])
def test_open_window_equals_no_window_bitwise(wave, tile):
    """A window open over the whole run takes the per-row checked code
    path in every kernel; it must equal the unchecked path bitwise."""
    _print("case: open window == none", wave, tile)
    p = Problem(wave, 128, 64, 10, 10.0, 7e-4, 200, seed=15, sources=[(50, 20)], rec_z=3)
    w = (np.zeros((p.nz, p.nx), np.float32), 1.0, 1.0)
    a, fa = _run(p, w, tile)
    b, fb = _run(p, None, tile)
    assert np.abs(b[0][0]).max() > 0
    for c in range(len(a[0])):
        assert np.array_equal(a[0][c], b[0][c])
    for n in fa[0]:
        assert np.array_equal(fa[0][n], fb[0][n]), n


def test_sh_block_shapes_under_a_window_run_the_two_pass_kernels():
    """A block shape set on a context with a first-arrival window runs tile
    0's two-pass kernels (the block step has no window): bitwise equal to
    tile 0 under the same window, and to the oracle."""
    _print("case: sh block shapes under a window")
    lib = _lib.load()
    blocks = [t for t in range(lib.fd23_num_tiles()) if lib.fd23_tile_kind(t) == _lib.TILE_BLOCK]
    p = Problem(_lib.WAVE_SH, 128, 64, 10, 10.0, 7e-4, 300, seed=13, kind=_lib.SRC_FORCE,
                sources=[(40, 12)], rec_z=2)
    w = _window(p)
    ref, rf = _run(p, w, 0)
    _check(p, ref, rf, w, 1e-5)
    for t in blocks:
        seis, fields = _run(p, w, t)
        assert np.array_equal(seis[0][0], ref[0][0]), t
        for n in fields[0]:
            assert np.array_equal(fields[0][n], rf[0][n]), (t, n)


def test_window_refusals():
    _print("case: punch refusals")
    p = Problem(_lib.WAVE_PSV, 128, 64, 10, 10.0, 7e-4, 10, seed=16, rec_z=3)
    t = np.zeros((p.nz, p.nx), np.float32)
    with p.context() as ctx:
        ctx.set_punch(t, 0.0, 1.0, p.dt)
        with pytest.raises(_lib.Fd23Error, match="first-arrival"):
            ctx.step_stress()
        with pytest.raises(_lib.Fd23Error, match="first-arrival"):
            ctx.step_velocity()
        ctx.set_punch(None)
        ctx.step_stress()  # staged stepping again without a window
        with pytest.raises(_lib.Fd23Error):
            ctx.set_punch(t, 0.0, 1.0, 0.0)  # dt
    with p.context(storage=_lib.STORE_F16) as ctx:
        with pytest.raises(_lib.Fd23Error, match="FP32 or FP64"):
            ctx.set_punch(t, 0.0, 1.0, p.dt)
    q = Problem(_lib.WAVE_SH, 100, 64, 10, 10.0, 7e-4, 10, seed=17, rec_z=3)
    with q.context() as ctx:
        with pytest.raises(_lib.Fd23Error, match="multiple"):
            ctx.set_punch(np.zeros((q.nz, q.nx), np.float32), 0.0, 1.0, q.dt)


def test_box_with_window_vs_oracle():
    """Hybrid box coupling (PSV/psv_clean_double/main.cpp:153-200) with a
    window: the ring add/sub and the residual grid run every step, the main
    passes only in open blocks (cu_step_stress/velocity take the window,
    the residual passes do not)."""
    import oracle as O
    from fd23_amd import host
    from fd23_amd.core import Context
    from helpers import layered_model, unpack_cpml

    _print("case: box + punch")
    nx, nz, npml, h, dt, nt = 256, 192, 20, 10.0, 8e-4, 300
    box = dict(x0=60, z0=50, bx=120, bz=90)
    med = dict(vp=3000.0, vs=1732.0, rho=2300.0)
    wave = dict(incident=0, angle_deg=20.0, f0=12.0, t0=1.5 / 12.0, amp=1.0)
    vp, vs, den = layered_model(nx, nz, 78, vmin=2000.0, vmax=4500.0)
    mat, _ = host.psv_material(vp, vs, den, h, dt)
    packed = host.cpml_pack(_lib.WAVE_PSV, nx, nz, npml, dt, 1e-6, float(vp.max()) / h, 12.0)
    # distance to the box (0 inside) at the fastest speed
    z, x = np.mgrid[0:nz, 0:nx]
    dx = np.maximum(np.maximum(box["x0"] - x, x - (box["x0"] + box["bx"] - 1)), 0)
    dz = np.maximum(np.maximum(box["z0"] - z, z - (box["z0"] + box["bz"] - 1)), 0)
    times = (np.hypot(dx, dz) * h / float(vp.max())).astype(np.float32)
    before, after = 0.01, 0.15
    rx = np.arange(2, nx - 2, 3, dtype=np.int32)
    rz = np.full(rx.size, 25, np.int32)

    cp = unpack_cpml(packed, _lib.WAVE_PSV, nx, nz, np.float32)
    sim = O.PSV(nx, nz, npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in mat])), cp)
    ob = O.PlaneWaveBox(sim, box["x0"], box["z0"], box["bx"], box["bz"], wave["incident"], wave["angle_deg"],
                        wave["f0"], wave["t0"], wave["amp"], med["vp"], med["vs"], med["rho"], h, dt)
    sim.set_window(times, before, after, dt)
    first, last = sim.window
    assert first.max() > 0 and last.min() < nt  # blocks open late and close early
    u, w = O.run_psv_box(sim, ob, nt, rx, rz)

    ctx = Context(_lib.WAVE_PSV, nx, nz, npml, max_frames=nt, max_rec=rx.size)
    try:
        ctx.set_material(mat)
        ctx.set_cpml(packed)
        for c in range(2):
            ctx.set_receivers(0, c, rx, rz)
        ctx.box_configure(box["x0"], box["z0"], box["bx"], box["bz"], wave["incident"], wave["angle_deg"],
                          wave["f0"], wave["t0"], wave["amp"], med["vp"], med["vs"], med["rho"], h, dt)
        ctx.set_punch(times, before, after, dt)
        ctx.run(0, nt)
        ctx.sync()
        errs = {"seis U": rel(ctx.seismograms(0, 0, nt), u), "seis W": rel(ctx.seismograms(0, 1, nt), w)}
        res = ctx.box_residual()
        # against the residual grid's scale of the same kind, as the box
        # tests measure (the global grid holds total minus incident inside)
        kinds = {"U": ("U", "W"), "W": ("U", "W"), "Txx": ("Txx", "Tzz", "Txz"),
                 "Tzz": ("Txx", "Tzz", "Txz"), "Txz": ("Txx", "Tzz", "Txz")}
        for name in ("U", "W", "Txx", "Tzz", "Txz"):
            scale = max(np.linalg.norm(ob.r[k]) for k in kinds[name])
            errs["field " + name] = float(np.linalg.norm(ctx.field(name).astype(np.float64) - getattr(sim, name))
                                          / scale)
            errs["resid " + name] = float(np.linalg.norm(res[name].astype(np.float64) - ob.r[name]) / scale)
    finally:
        ctx.close()
    _print(errs)
    assert np.abs(u).max() > 0
    bad = {k: v for k, v in errs.items() if not v < 1e-5}
    assert not bad, bad


def test_autotune_under_a_window_picks_a_shape_that_runs():
    """Under a window split shapes run unsplit and fused shapes run as tile
    0, so the tuner must not choose them (it would time one shape and run
    another)."""
    _print("case: autotune with a window")
    p = _psv_problem()
    with p.context() as ctx:
        ctx.set_punch(*_window(p), p.dt)
        tile, _ = ctx.autotune(2)
        assert ctx.tile == tile
        assert ctx.tile_kind() == _lib.TILE_TILED and not ctx.tile_split()
