"""The oracle's first-arrival window ("punch", SH/slide/cu_slide.cu:7-61,
SH/fd/cu_fd.cu:18-25) on the CPU.

- the per-block step window against a direct per-block loop in float32
  scalars (the reference's arithmetic: block min/max, minus before / plus
  after, divided by dt, truncated toward zero), with a known answer for a
  constant block and for a negative quotient;
- one pass with a mixed window from a random state: cells of closed blocks
  (field and C-PML memory variables) keep their old values, cells of open
  blocks equal the same pass without a window;
- SH: the surface imaging of Tz row 0 runs only in open columns;
- a window open over the whole run equals no window, bitwise; a window
  closed over the whole run leaves only the injected source values.
"""
import numpy as np
import pytest

import oracle as O
from helpers import layered_model


def _window_loop(times, before, after, dt):
    nz, nx = times.shape
    f32 = np.float32
    first = np.zeros((nz // 16, nx // 32), np.int32)
    last = np.zeros_like(first)
    for k in range(nz // 16):
        for i in range(nx // 32):
            lo, hi = f32(1e28), f32(-1e28)
            for kk in range(16):
                for ii in range(32):
                    t = f32(times[k * 16 + kk, i * 32 + ii])
                    lo, hi = min(lo, t), max(hi, t)
            first[k, i] = int((lo - f32(before)) / f32(dt))
            last[k, i] = int((hi + f32(after)) / f32(dt))
    return first, last


def test_window_matches_per_block_loop():
    rng = np.random.default_rng(3)
    t = rng.uniform(-0.05, 0.4, (48, 96)).astype(np.float32)
    first, last = O.punch_window(t, 0.013, 0.021, 7e-4)
    f2, l2 = _window_loop(t, 0.013, 0.021, 7e-4)
    assert np.array_equal(first, f2) and np.array_equal(last, l2)
    # constant block: (0.1 - 0.02) / 0.01 = 8 in float32 -> 7 or 8; the
    # quotient below zero truncates toward zero
    c = np.full((16, 32), 0.1, np.float32)
    f, l = O.punch_window(c, 0.02, 0.05, 0.01)
    assert f[0, 0] == int((np.float32(0.1) - np.float32(0.02)) / np.float32(0.01))
    assert l[0, 0] == int((np.float32(0.1) + np.float32(0.05)) / np.float32(0.01))
    f, _ = O.punch_window(np.full((16, 32), 0.005, np.float32), 0.01, 0.0, 0.01)
    assert f[0, 0] == 0  # -0.5 truncates to 0, not -1


def test_mask_is_the_open_blocks():
    first = np.array([[0, 4], [-1, 9]], np.int32)
    last = np.array([[4, 9], [10, 12]], np.int32)
    m = O.punch_mask(first, last, 4)  # open: first < 4 < last, both strict
    assert m.shape == (32, 64)
    want = np.array([[0, 0], [1, 0]], bool)  # (0,0): it == last; (0,1): it == first; (1,1): it < first
    for k in range(2):
        for i in range(2):
            assert (m[k * 16:(k + 1) * 16, i * 32:(i + 1) * 32] == want[k, i]).all()


def _psv(nx, nz, npml, top, seed):
    vp, vs, den = layered_model(nx, nz, seed)
    h, dt = 10.0, 7e-4
    mat, _ = O.psv_material(vp, vs, den, h, dt)
    cp = O.psv_cpml(nx, nz, npml, dt, 1e-6, float(vp.max()) / h, 12.0)
    return O.PSV(nx, nz, npml, mat, cp, topabsorb=top)


def _sh(nx, nz, npml, top, seed):
    vp, vs, den = layered_model(nx, nz, seed)
    h, dt = 10.0, 7e-4
    mat, _ = O.sh_material(vs, den, h, dt)
    cp = O.sh_cpml(nx, nz, npml, dt, 1e-6, float(vp.max()) / h, 12.0)
    return O.SH(nx, nz, npml, mat, cp, topabsorb=top)


def _psi_cells(nx, nz, npml, top, xdir):
    """Cell (iz, ix) of every entry of a psi array (the passes' layout: x
    strips [iz][2*npml], z strips [2*npml][nx]); iz = -1 where none maps."""
    if xdir:
        iz, j = np.divmod(np.arange(2 * npml * nz), 2 * npml)
        return iz, np.where(j < npml, j, nx - 1 - (j - npml))
    r, ix = np.divmod(np.arange(2 * npml * nx), nx)
    return np.where(r < npml, r if top else -1, nz - 1 - (r - npml)), ix


FIELDS = {"psv": ("U", "W", "Txx", "Txz", "Tzz"), "sh": ("V", "Tx", "Tz")}
# what each pass writes: fields, then psi arrays
WRITES = {("psv", "stress"): (("Txx", "Tzz", "Txz"), ("U_x", "W_x", "U_z", "W_z")),
          ("psv", "velocity"): (("U", "W"), ("Txx_x", "Txz_x", "Txz_z", "Tzz_z")),
          ("sh", "stress"): (("Tx", "Tz"), ("V_x", "V_z")),
          ("sh", "velocity"): (("V",), ("Tx_x", "Tz_z"))}
XPSI = set(O.PSV_CPML_X) | {"Tx_x", "V_x"}


@pytest.mark.parametrize("wave", ["psv", "sh"])
@pytest.mark.parametrize("top", [False, True])
@pytest.mark.parametrize("pas", ["stress", "velocity"])
def test_closed_blocks_keep_their_values(wave, top, pas):
    nx, nz, npml = 128, 64, 10
    rng = np.random.default_rng(5)
    make = _psv if wave == "psv" else _sh
    a, b = make(nx, nz, npml, top, 2), make(nx, nz, npml, top, 2)
    for n in FIELDS[wave]:
        v = rng.standard_normal((nz, nx)).astype(np.float32)
        getattr(a, n)[...] = v
        getattr(b, n)[...] = v
    for n in a.psi:
        a.psi[n][...] = b.psi[n][...] = rng.standard_normal(a.psi[n].shape).astype(np.float32)
    s0 = {n: getattr(a, n).copy() for n in FIELDS[wave]}
    p0 = {n: v.copy() for n, v in a.psi.items()}
    # a checkerboard of open blocks at step 7, the top block row closed in
    # the left half and open in the right half (the SH imaging case)
    first = np.zeros((nz // 16, nx // 32), np.int32)
    last = np.full_like(first, 20)
    first[(np.add.outer(np.arange(nz // 16), np.arange(nx // 32)) % 2) == 1] = 9
    first[0, :2], first[0, 2:] = 9, 0
    b.window = (first, last)
    op = O.punch_mask(first, last, 7).astype(bool)
    assert 0 < op.sum() < op.size
    getattr(a, pas)()
    getattr(b, pas)(7)
    fields, psis = WRITES[(wave, pas)]
    for n in fields:
        fa, fb, f0 = getattr(a, n), getattr(b, n), s0[n]
        assert np.array_equal(fb[op], fa[op]), n
        assert np.array_equal(fb[~op], f0[~op]), n
        assert not np.array_equal(fa[~op], f0[~op]), n  # the window made a difference
    if wave == "sh" and pas == "velocity":
        # Tz row 0 is imaged (= -Tz row 1) in open columns only
        o = op[0]
        assert np.array_equal(b.Tz[0][o], -b.Tz[1][o])
        assert np.array_equal(b.Tz[0][~o], s0["Tz"][0][~o])
        assert np.array_equal(b.Tz[1:], s0["Tz"][1:])
    for n in psis:
        iz, ix = _psi_cells(nx, nz, npml, top, n in XPSI)
        live = iz >= 0
        m = np.zeros(iz.shape, bool)
        m[live] = op[iz[live], ix[live]]
        assert np.array_equal(b.psi[n][m], a.psi[n][m]), n
        assert np.array_equal(b.psi[n][~m], p0[n][~m]), n
        assert not np.array_equal(a.psi[n][~m & live], p0[n][~m & live]), n


def test_window_always_open_equals_no_window():
    nx, nz, npml, nt = 96, 48, 8, 60
    a, b = _psv(nx, nz, npml, False, 4), _psv(nx, nz, npml, False, 4)
    b.set_window(np.zeros((nz, nx), np.float32), 1.0, 1.0, 7e-4)  # first < 0, last > nt
    assert (b.window[0] < 0).all() and (b.window[1] > nt).all()
    src = O.ricker(nt, 7e-4, 12.0).astype(np.float32)
    M = np.array([1.0, 0.5, -0.3], np.float32)
    ua, _ = O.run_psv(a, nt, src, M, 40, 20, rx=np.arange(1, nx, 3), rz=np.full(32, 5))
    ub, _ = O.run_psv(b, nt, src, M, 40, 20, rx=np.arange(1, nx, 3), rz=np.full(32, 5))
    assert np.abs(ua).max() > 0
    assert np.array_equal(ua, ub)
    for n in FIELDS["psv"]:
        assert np.array_equal(getattr(a, n), getattr(b, n)), n


def test_window_always_closed_leaves_the_source_only():
    nx, nz, npml, nt = 64, 32, 6, 30
    a, b = _sh(nx, nz, npml, False, 6), _sh(nx, nz, npml, False, 6)
    b.set_window(np.zeros((nz, nx), np.float32), 0.0, -1.0, 1e-3)  # last < 0 < first + 1
    src = O.ricker(nt, 1e-3, 20.0).astype(np.float32)
    M = np.array([1.0, 0.5], np.float32)
    for it in range(nt):
        b.stress(it)
        b.inject(0, it, src, M, 30, 12)
        b.velocity(it)
        a.inject(0, it, src, M, 30, 12)  # injection alone, no passes
    for n in FIELDS["sh"]:
        assert np.array_equal(getattr(a, n), getattr(b, n)), n
    assert np.abs(b.Tx).max() > 0


def test_bad_window_is_refused():
    sim = _sh(64, 32, 6, False, 1)
    sim.set_window(np.zeros((32, 64), np.float32), 0.0, 0.0, 1e-3)
    with pytest.raises(ValueError):
        sim.stress()  # no step index
    with pytest.raises(ValueError):
        sim.set_window(np.zeros((32, 32), np.float32), 0.0, 0.0, 1e-3)
    with pytest.raises(ValueError):
        _sh(64, 30, 6, False, 1).set_window(np.zeros((30, 64), np.float32), 0.0, 0.0, 1e-3)


def _window_loop_f64(times, before, after, dt):
    # the FP64 tree's SLIDE::SLIDE in Python floats (IEEE double)
    nz, nx = times.shape
    first = np.zeros((nz // 16, nx // 32), np.int32)
    last = np.zeros_like(first)
    for k in range(nz // 16):
        for i in range(nx // 32):
            blk = [float(v) for v in times[k * 16:(k + 1) * 16, i * 32:(i + 1) * 32].ravel()]
            first[k, i] = int((min(blk) - before) / dt)
            last[k, i] = int((max(blk) + after) / dt)
    return first, last


def test_window_tables_float_and_double_trees():
    """The library's host window tables (fd23_punch_blocks[_f64]) equal the
    oracle and a direct loop, in float for the FP32 trees and in double for
    the FP64 tree, on travel times chosen so that the two precisions
    truncate differently: 0.3/0.1 is 2.9999999999999996 in double (2) and
    exactly 3 in float32 (3)."""
    from fd23_amd import host

    rng = np.random.default_rng(11)
    t = rng.uniform(0.0, 0.6, (32, 64))
    # block (0,0): minimum exactly 0.3 -> first differs between precisions
    t[0:16, 0:32] = np.maximum(t[0:16, 0:32], 0.3)
    t[3, 5] = 0.3
    # block (1,1): maximum 0.7 with after 0 -> 0.7/0.1 = 6.999999999999999 (6) in double, 7 in float32
    t[16:32, 32:64] = np.minimum(t[16:32, 32:64], 0.7)
    t[20, 40] = 0.7
    before, after, dt = 0.0, 0.0, 0.1
    f32f, f32l = host.punch_blocks(t, before, after, dt)
    f64f, f64l = host.punch_blocks(t, before, after, dt, f64=True)
    of, ol = O.punch_window(t, before, after, dt)
    df, dl = O.punch_window(t, before, after, dt, dtype=np.float64)
    lf, ll = _window_loop_f64(t, before, after, dt)
    assert np.array_equal(f32f, of) and np.array_equal(f32l, ol)
    assert np.array_equal(f64f, df) and np.array_equal(f64l, dl)
    assert np.array_equal(f64f, lf) and np.array_equal(f64l, ll)
    assert f64f[0, 0] == 2 and f32f[0, 0] == 3
    assert (f64l != f32l).any() or (f64f != f32f).sum() >= 1
    # float32 tables of the float32 loop too
    g, h = _window_loop(t.astype(np.float32), before, after, dt)
    assert np.array_equal(f32f, g) and np.array_equal(f32l, h)


def test_window_tables_reject_bad_grids():
    from fd23_amd import _lib, host

    with pytest.raises(_lib.Fd23Error):
        host.punch_blocks(np.zeros((16, 48)), 0.0, 0.0, 0.1)
    with pytest.raises(_lib.Fd23Error):
        host.punch_blocks(np.zeros((16, 32)), 0.0, 0.0, 0.0, f64=True)
