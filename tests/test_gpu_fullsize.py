"""The benchmark configurations at full size on the GPU, checked through
properties that hold at any size (the oracle is too slow for these grids;
the parity tests pin the same kernels to it on smaller grids):

- accuracy: the 2000x1000 P-SV configuration over its 5000 steps, FP32
  storage against the FP64-storage run of the same inputs (the FP64 path
  equals the FP64 oracle to 1e-10 in test_gpu_parity): seismograms within
  1e-5 relative L2;
- linearity, FP64 storage: the same run with the source doubled gives
  twice the seismograms and fields, to 1e-15 of the peak over the first
  300 steps and to 1e-12 over all 5000.  It cannot be exact: the cells
  ahead of the wave front fall through the sub-normal range (values decay
  by orders of magnitude per cell there), where numbers are rounded to a
  fixed step and doubling is not exact; after the first differing bit the
  two runs round independently and drift apart at roundoff level.  FP32
  storage reaches that range within a few steps (its runs differ by ~1e-5
  relative after 5000 steps, as the CPU oracle's do), so it is not tested
  for linearity;
- mirror symmetry: a 2001x1000 model that varies in z only, with an
  explosion at the centre column, gives W, Txx, Tzz symmetric and U, Txz
  antisymmetric about that column, with each field's half-cell staggering,
  bitwise;
- batch == single: shots of the 8-shot 4096x2048 configuration run in one
  context equal the same shots run alone, bitwise;
- the 400x200 SH configuration's full run and the first and last shot of
  the 64-shot 1024x512 batch (300 steps) against the FP32 oracle (1e-5).
"""
import numpy as np
import pytest

from fd23_amd import _lib, workloads as W
from helpers import rel

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _run(wl, nt, stf_scale=1.0, shots=None):
    if stf_scale != 1.0:
        wl.stf = wl.stf * stf_scale
    ctx = wl.context(max_frames=nt, nshots=shots)
    try:
        ctx.run(0, nt, 1)
        ctx.sync()
        seis = np.stack([ctx.seismograms(s, c, nt).astype(np.float64) for s in range(ctx.nshots)
                         for c in range(2)])
        fields = {n: ctx.field(n).astype(np.float64) for n in ("U", "W", "Txx", "Tzz", "Txz")}
        states = [ctx.get_state(s) for s in range(ctx.nshots)]
    finally:
        ctx.close()
    if stf_scale != 1.0:
        wl.stf = wl.stf / stf_scale
    return seis, fields, states


def test_c2_fp32_against_fp64_full_run():
    _print("case: 2000x1000 fp32 vs fp64, 5000 steps")
    wl = W.config_psv_2000x1000().build()
    a, _, _ = _run(wl, wl.nt)
    w64 = W.config_psv_2000x1000()
    w64.storage = _lib.STORE_F64
    w64.build()
    b, _, _ = _run(w64, w64.nt)
    e = rel(a, b)
    _print("fp32 vs fp64 seismograms", e)
    assert np.abs(b).max() > 0
    assert e < 1e-5


def test_c2_linearity_fp64():
    _print("case: 2000x1000 linearity, FP64 storage")
    wl = W.config_psv_2000x1000()
    wl.storage = _lib.STORE_F64
    wl.build()

    def dev(x, y):  # largest deviation of y from 2x, relative to the largest |2x|
        return float(np.abs(y - 2.0 * x).max() / np.abs(2.0 * x).max())

    for nt, tol in ((300, 1e-15), (wl.nt, 1e-12)):
        a, fa, _ = _run(wl, nt)
        b, fb, _ = _run(wl, nt, stf_scale=2.0)
        assert np.abs(a).max() > 0
        d = max([dev(a, b)] + [dev(fa[n], fb[n]) for n in fa])
        _print(nt, "steps: largest deviation from 2x, relative to the peak", d)
        assert d < tol


def test_mirror_symmetry_full_size():
    _print("case: 2001x1000 mirror symmetry, 1500 steps")
    nx, nz, xs = 2001, 1000, 1000
    wl = W.config_psv_2000x1000()
    wl.nx, wl.sources, wl.nt = nx, [(xs, 10)], 1500
    vp, vs, den = W.layered_grf_model(nx, nz, wl.dx, wl.seed)
    # the layering of the benchmark model, constant along x
    vp, vs, den = (np.ascontiguousarray(np.repeat(a[:, :1], nx, axis=1)) for a in (vp, vs, den))
    wl.build(vp, vs, den)
    _, f, _ = _run(wl, wl.nt)
    assert np.abs(f["Txx"]).max() > 0
    cols = np.arange(1, nx - 1)
    for name, sign, mirror in (("Txx", 1, lambda x: 2 * xs - x), ("Tzz", 1, lambda x: 2 * xs - x),
                               ("W", 1, lambda x: 2 * xs - x), ("U", -1, lambda x: 2 * xs - 1 - x),
                               ("Txz", -1, lambda x: 2 * xs - 1 - x)):
        m = mirror(cols)
        ok = (m >= 0) & (m < nx)
        assert np.array_equal(f[name][:, cols[ok]], sign * f[name][:, m[ok]]), name


def test_c4_batch_equals_single_shots():
    _print("case: 4096x2048 x8 batch vs single, 200 steps")
    nt = 200
    wl = W.config_psv_precision(_lib.STORE_F32).build()
    seis, _, states = _run(wl, nt)
    for k in (0, 5):
        one = W.config_psv_precision(_lib.STORE_F32)
        one.sources = [wl.sources[k]]
        one.build()
        s1, _, st1 = _run(one, nt)
        _print("shot", k, st1[0] == states[k])
        assert st1[0] == states[k]
        assert np.array_equal(s1, seis[2 * k:2 * k + 2])


def _oracle_for(wl, shot, nt):
    """The oracle's run of one shot of a workload (same material, C-PML,
    wavelet and moment tensor as the context gets), FP32."""
    import oracle as O
    from helpers import unpack_cpml

    cp = unpack_cpml(wl.packed, wl.wave, wl.nx, wl.nz, np.float32)
    xs, zs = wl.sources[shot]
    rx, rz = wl.receivers()
    stf, M = wl.stf.astype(np.float32), wl.M.astype(np.float32)
    if wl.wave == _lib.WAVE_PSV:
        sim = O.PSV(wl.nx, wl.nz, wl.npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), wl.mat)), cp, wl.topabsorb)
        kind = {_lib.SRC_LINE: 0, _lib.SRC_P: 1}[wl.kind]
        u, w = O.run_psv(sim, nt, stf, M, xs, zs, kind=kind, rx=rx, rz=rz)
        return sim, [u, w]
    sim = O.SH(wl.nx, wl.nz, wl.npml, dict(zip(("BV", "MUX", "MUZ"), wl.mat)), cp, wl.topabsorb)
    v = O.run_sh(sim, nt, stf, M, xs, zs, kind=_lib.SRC_FORCE, rx=rx, rz=rz)
    return sim, [v]


def test_c1_sh_full_run_vs_oracle():
    """The 400x200 SH configuration over its 2000 steps against the FP32
    oracle, on the two-pass tiles (tile 0) and on every block-step shape:
    seismograms within 1e-5 relative L2; the fields at step 700 (the wave
    inside the grid) within 1e-5 relative L2, and at the end, when the
    absorbed wave has left small residuals, within 1e-5 of the step-700
    peak.  The receivers at z=2 lie in the frozen rows of the absorbing top
    (C-PML profiles are zero for the first 4 rows, cpml.cpp:54-64): rows
    0-3 only couple along x and stay zero, so both sides record zeros
    there."""
    import oracle as O
    from helpers import unpack_cpml

    _print("case: sh 400x200 full run vs oracle")
    wl = W.config_sh_400x200().build()
    names = ("V", "Tx", "Tz")
    mid = 700
    cp = unpack_cpml(wl.packed, wl.wave, wl.nx, wl.nz, np.float32)
    sim = O.SH(wl.nx, wl.nz, wl.npml, dict(zip(("BV", "MUX", "MUZ"), wl.mat)), cp, wl.topabsorb)
    rx, rz = wl.receivers()
    stf, M = wl.stf.astype(np.float32), wl.M.astype(np.float32)
    xs, zs = wl.sources[0]
    v1 = O.run_sh(sim, mid, stf, M, xs, zs, kind=_lib.SRC_FORCE, rx=rx, rz=rz)
    o_mid = {n: getattr(sim, n).copy() for n in names}
    v2 = O.run_sh(sim, wl.nt - mid, stf, M, xs, zs, kind=_lib.SRC_FORCE, rx=rx, rz=rz, it0=mid)
    o_seis = np.concatenate([v1, v2])
    lib = _lib.load()
    tiles = [0] + [t for t in range(lib.fd23_num_tiles()) if lib.fd23_tile_kind(t) == _lib.TILE_BLOCK]
    for tile in tiles:
        ctx = wl.context()
        try:
            ctx.tile = tile
            ctx.run(0, mid, 1)
            f_mid = {n: ctx.field(n) for n in names}
            ctx.run(mid, wl.nt - mid, 1)
            ctx.sync()
            g = ctx.seismograms(0, 0, wl.nt)
            f_end = {n: ctx.field(n) for n in names}
        finally:
            ctx.close()
        e = rel(g, o_seis)
        _print("tile", tile, "seis", e)
        assert e < 1e-5
        for n in names:
            ref, peak = o_mid[n], float(np.abs(o_mid[n]).max())
            e_mid = rel(f_mid[n], ref)
            e_end = float(np.abs(f_end[n].astype(np.float64) - getattr(sim, n)).max() / peak)
            _print(n, "step", mid, e_mid, "end, of the step-%d peak" % mid, e_end)
            assert peak > 0 and e_mid < 1e-5 and e_end < 1e-5
            assert not ref[:4].any() and not f_mid[n][:4].any()  # the frozen top rows


def test_c3_batch_first_and_last_shot_vs_oracle():
    """The 1024x512 x 64-shot batch (300 steps): the first and the last
    shot of the one launch against the FP32 oracle run of that shot."""
    _print("case: 1024x512 x64 shots 0 and 63 vs oracle, 300 steps")
    nt = 300
    wl = W.config_psv_batched(64).build()
    ctx = wl.context(max_frames=nt)
    try:
        ctx.run(0, nt, 1)
        ctx.sync()
        got = {s: [ctx.seismograms(s, c, nt) for c in range(2)] for s in (0, 63)}
        fld = {s: ctx.field("Txx", s) for s in (0, 63)}
    finally:
        ctx.close()
    for s in (0, 63):
        sim, ref = _oracle_for(wl, s, nt)
        errs = [rel(got[s][c], ref[c]) for c in range(2)] + [rel(fld[s], sim.Txx)]
        _print("shot", s, errs)
        assert np.abs(ref[0]).max() > 0
        assert max(errs) < 1e-5
