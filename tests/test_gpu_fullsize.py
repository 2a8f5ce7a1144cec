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
  context equal the same shots run alone, bitwise.
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
