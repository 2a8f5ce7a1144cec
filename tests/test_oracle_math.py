"""Checks of the oracle that the mathematics fixes (CPU).

- the 8th-order staggered operator differentiates polynomials up to degree 8
  exactly at the half-cell position (a transposed tap or wrong sign fails)
- mirror symmetry: an explosive P-SV source and an SH point force in a
  homogeneous model give fields that are exactly (anti)symmetric about the
  source column, with the half-cell staggering of each field
- linearity: scaling the source by a power of two scales every field by it
  exactly; a zero source leaves every field at zero
"""
import numpy as np
import pytest

import oracle as O


def _unit_psv(nx, nz, dtype=np.float64):
    """Material with LAM=0, MU=1/2 (gam = 1), MUA=1, BU=BW=1, no C-PML."""
    z = lambda v: np.full((nz, nx), v, dtype)
    mat = {"LAM": z(0.0), "MU": z(0.5), "MUA": z(1.0), "BU": z(1.0), "BW": z(1.0)}
    cp = {k: (np.zeros(n, dtype), np.zeros(n, dtype), np.zeros(n, dtype))
          for k, n in [(k, nx) for k in O.PSV_CPML_X] + [(k, nz) for k in O.PSV_CPML_Z]}
    return O.PSV(nx, nz, 0, mat, cp, dtype=dtype)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 6, 7, 8])
def test_stencil_exact_on_polynomials(deg):
    nx, nz = 40, 40
    sim = _unit_psv(nx, nz)
    x = np.arange(nx, dtype=np.float64)
    x0 = 17.3
    sim.U[:] = ((x - x0) / 10.0) ** deg            # U depends on x only
    sim.W[:] = (((np.arange(nz) - x0) / 10.0) ** deg)[:, None]  # W on z only
    sim.stress()
    # Txx = gam*ux + lam*wz = ux (d/dx at x-1/2);  Tzz = lam*ux + gam*wz = wz
    # (d/dz at z+1/2); interior cells, full 8th order
    ux_true = deg * ((x - 0.5 - x0) / 10.0) ** (deg - 1) / 10.0
    wz_true = deg * ((np.arange(nz) + 0.5 - x0) / 10.0) ** (deg - 1) / 10.0
    sl = slice(5, 35)
    np.testing.assert_allclose(sim.Txx[20, sl], ux_true[sl], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(sim.Tzz[sl, 20], wz_true[sl], rtol=1e-10, atol=1e-12)


def test_stencil_not_exact_on_degree_9():
    nx = 40
    sim = _unit_psv(nx, nx)
    x = np.arange(nx, dtype=np.float64)
    sim.U[:] = ((x - 17.3) / 10.0) ** 9
    sim.stress()
    true = 9 * ((x - 0.5 - 17.3) / 10.0) ** 8 / 10.0
    assert np.abs(sim.Txx[20, 5:35] - true[5:35]).max() > 1e-9


def _homog_psv(nx, nz, npml, dtype):
    vp = np.full((nz, nx), 3000.0, dtype)
    mat, _ = O.psv_material(vp, vp / 1.732, vp * 0 + 2500.0, 10.0, 1e-3, dtype)
    cp = O.psv_cpml(nx, nz, npml, 1e-3, 1e-6, 300.0, 15.0, dtype)
    return O.PSV(nx, nz, npml, mat, cp, dtype=dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_psv_mirror_symmetry(dtype):
    nx, nz, xs = 121, 80, 60
    sim = _homog_psv(nx, nz, 10, dtype)
    stf = O.ricker(60, 1e-3, 25.0).astype(dtype)
    M = O.psv_moment(2, 0, 0, 0, 0, 10.0, 1e-3, dtype)  # explosion
    O.run_psv(sim, 60, stf, M, xs, 20)
    assert np.abs(sim.Txx).max() > 0
    # integer-grid fields are symmetric about xs, half-grid ones antisymmetric
    # about xs - 1/2 + ... : U(x+1/2) = -U(2 xs - x - 1/2) -> column 2xs-1-x
    for f, sign, mirror in ((sim.Txx, 1, lambda x: 2 * xs - x), (sim.Tzz, 1, lambda x: 2 * xs - x),
                            (sim.W, 1, lambda x: 2 * xs - x), (sim.U, -1, lambda x: 2 * xs - 1 - x),
                            (sim.Txz, -1, lambda x: 2 * xs - 1 - x)):
        cols = np.arange(30, 91)
        assert np.array_equal(f[:, cols], sign * f[:, mirror(cols)])


def test_psv_symmetry_breaks_with_a_sign_error():
    """Flip one tap sign of the source pattern: the symmetry test catches it."""
    nx, nz, xs = 121, 80, 60
    sim = _homog_psv(nx, nz, 10, np.float64)
    stf = O.ricker(60, 1e-3, 25.0)
    M = O.psv_moment(2, 0, 0, 0, 0, 10.0, 1e-3, np.float64)
    for it in range(60):
        sim.stress()
        sim.inject(2, it, stf, M, xs, 20)
        sim.Txx[20, xs + 1] -= 1e-3 * stf[it] * M[0]  # planted asymmetry
        sim.velocity()
    cols = np.arange(30, 91)
    assert not np.array_equal(sim.Txx[:, cols], sim.Txx[:, 2 * xs - cols])


def test_sh_mirror_symmetry():
    nx, nz, xs, npml = 121, 80, 60, 10
    vs = np.full((nz, nx), 2000.0, np.float64)
    mat, _ = O.sh_material(vs, vs * 0 + 2200.0, 10.0, 1e-3, np.float64)
    cp = O.sh_cpml(nx, nz, npml, 1e-3, 1e-6, 200.0, 15.0, np.float64)
    sim = O.SH(nx, nz, npml, mat, cp, dtype=np.float64)
    stf = O.ricker(80, 1e-3, 25.0)
    M = O.sh_moment(0, 0, 90, 0, 0, 10.0, 1e-3, np.float64)
    O.run_sh(sim, 80, stf, M, xs, 20, kind=2)  # point force on V
    cols = np.arange(30, 91)
    assert np.abs(sim.V).max() > 0
    assert np.array_equal(sim.V[:, cols], sim.V[:, 2 * xs - cols])
    assert np.array_equal(sim.Tz[:, cols], sim.Tz[:, 2 * xs - cols])
    assert np.array_equal(sim.Tx[:, cols], -sim.Tx[:, 2 * xs - 1 - cols])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_linearity_and_zero_source(dtype):
    nx, nz = 90, 70
    rng = np.random.default_rng(4)
    vp = (3000 + 200 * rng.standard_normal((nz, nx))).astype(dtype)
    mat, _ = O.psv_material(vp, vp / 1.8, vp * 0 + 2400.0, 10.0, 1e-3, dtype)
    cp = O.psv_cpml(nx, nz, 8, 1e-3, 1e-6, 320.0, 15.0, dtype)
    stf = O.ricker(90, 1e-3, 20.0).astype(dtype)
    M = O.psv_moment(0, 20, 50, 70, 10, 10.0, 1e-3, dtype)
    a, b, z = (O.PSV(nx, nz, 8, mat, cp, dtype=dtype) for _ in range(3))
    O.run_psv(a, 90, stf, M, 40, 12)
    O.run_psv(b, 90, stf * 4, M, 40, 12)
    O.run_psv(z, 90, stf * 0, M, 40, 12)
    for n in ("U", "W", "Txx", "Tzz", "Txz"):
        if dtype == np.float64:
            assert np.array_equal(getattr(b, n), 4 * getattr(a, n))
        else:
            # FP32: exact until the stencil precursors ahead of the wave
            # underflow to subnormals, whose rounding is not scale-invariant
            x, y = getattr(b, n).astype(np.float64), 4 * getattr(a, n).astype(np.float64)
            assert np.linalg.norm(x - y) <= 1e-6 * np.linalg.norm(y)
        assert not np.any(getattr(z, n))


def _unit_sh(nx, nz, dtype=np.float64):
    """SH material MUX = MUZ = BV = 1, no C-PML."""
    z = lambda v: np.full((nz, nx), v, dtype)
    mat = {"BV": z(1.0), "MUX": z(1.0), "MUZ": z(1.0)}
    cp = {k: (np.zeros(n, dtype), np.zeros(n, dtype), np.zeros(n, dtype))
          for k, n in (("Tx_x", nx), ("V_x", nx), ("Tz_z", nz), ("V_z", nz))}
    return O.SH(nx, nz, 0, mat, cp, dtype=dtype)


@pytest.mark.parametrize("deg", [1, 3, 5, 8])
def test_sh_stencils_exact_on_polynomials(deg):
    """SH stress: Tx += MUX Dx+ V (at x+1/2), Tz += MUZ Dz- V (at z-1/2);
    SH velocity: V += BV (Dx- Tx + Dz+ Tz) -- exact up to degree 8."""
    n, x0 = 40, 17.3
    a = np.arange(n, dtype=np.float64)
    p = lambda t: ((t - x0) / 10.0) ** deg
    dp_ = lambda t: deg * ((t - x0) / 10.0) ** (deg - 1) / 10.0
    sim = _unit_sh(n, n)
    sim.V[:] = p(a)[None, :] + p(a)[:, None]
    sim.stress()
    sl = slice(5, 35)
    np.testing.assert_allclose(sim.Tx[20, sl], dp_(a + 0.5)[sl], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(sim.Tz[sl, 20], dp_(a - 0.5)[sl], rtol=1e-10, atol=1e-12)
    sim = _unit_sh(n, n)
    sim.Tx[:] = p(a)[None, :]
    sim.Tz[:] = p(a)[:, None]
    sim.velocity()
    # Dx- Tx at x - 1/2 plus Dz+ Tz at z + 1/2
    np.testing.assert_allclose(sim.V[20, sl], dp_(a - 0.5)[sl] + dp_(20 + 0.5), rtol=1e-10, atol=1e-12)


def test_sh_stress_skips_edge_cells_and_images_the_top_row():
    """SH/fd/cu_fd.cu:140 (no stress update at ix == 0, ix >= nx-1,
    iz >= nz-1) and :226 (Tz row 0 = -Tz row 1 before the velocity pass)."""
    n = 30
    sim = _unit_sh(n, n)
    rng = np.random.default_rng(3)
    sim.V[:] = rng.standard_normal((n, n))
    sim.stress()
    assert not sim.Tx[:, 0].any() and not sim.Tx[:, -1].any() and not sim.Tx[-1, :].any()
    assert not sim.Tz[:, 0].any() and not sim.Tz[-1, :].any()
    assert sim.Tx[1:-1, 1:-1].all()
    tz1 = sim.Tz[1].copy()
    sim.velocity()
    assert np.array_equal(sim.Tz[0], -tz1)
