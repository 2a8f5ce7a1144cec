"""The oracle against the reference's own fixtures (CPU).

  PSV/clean_nobox/TEST/tt/state000000, state000001: 1 and 2 steps of the
    96x96 top-absorbing run of tt/t1 (homo.mdl BACK 6.8 4.0 3.0)
  state000499 (500 steps) and state000999 (restart of t2 from 000499)
  U/W snapshots at it = 1, 100, 300, 500, 700, 900
  PSV/psv_clean_double/test/0_U, 0_W: FP64 2048x1024, 2001 steps,
    polygon model with a water layer (slow; bit-level agreement)
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref")
TT = os.path.join(GOLD, "psv_tt")


def _tt_sim():
    n, npml, h, dt = 96, 32, 0.3, 0.01
    vp, vs, den = O.mdl_raster(os.path.join(TT, "homo_6.8.mdl"), n, n, h)
    mat, stab = O.psv_material(vp, vs, den, h, dt)
    assert stab < 1
    cp = O.psv_cpml(n, n, npml, 0.005, 1e-11, 30.0, 2.0)
    return O.PSV(n, n, npml, mat, cp, topabsorb=True)


def _tt_src():
    return O.gaussian(-30), O.psv_moment(0, 0, 20, 90, 45, 0.3, 0.01)


def _state(name):
    return np.fromfile(os.path.join(TT, name), np.float32)


def test_state_step0_step1():
    sim = _tt_sim()
    src, M = _tt_src()
    O.run_psv(sim, 1, src, M, 33, 33)
    assert rel(np.frombuffer(sim.state_bytes(), np.float32), _state("state000000")) < 1e-7
    O.run_psv(sim, 1, src, M, 33, 33, it0=1)
    assert rel(np.frombuffer(sim.state_bytes(), np.float32), _state("state000001")) < 1e-8


def test_state_500_restart_and_snapshots():
    sim = _tt_sim()
    src, M = _tt_src()
    snaps = {}

    def cb(it, s):
        if it in (1, 100, 300):
            snaps[it] = (s.U.copy(), s.W.copy())

    O.run_psv(sim, 500, src, M, 33, 33, snap_cb=cb)
    assert rel(np.frombuffer(sim.state_bytes(), np.float32), _state("state000499")) < 2e-6
    for it, (u, w) in snaps.items():
        ru = np.fromfile(os.path.join(TT, "U%05d" % it), np.float32).reshape(96, 96)
        rw = np.fromfile(os.path.join(TT, "W%05d" % it), np.float32).reshape(96, 96)
        assert rel(u, ru) < 1e-5 and rel(w, rw) < 1e-5
    # restart from the reference's own checkpoint (tt/t2: restart=1 nt0=500)
    sim2 = _tt_sim()
    sim2.load_state(_state("state000499").tobytes())
    snaps2 = {}
    O.run_psv(sim2, 500, src, M, 33, 33, it0=500,
              snap_cb=lambda it, s: snaps2.__setitem__(it, (s.U.copy(), s.W.copy())) if it == 500 else None)
    assert rel(np.frombuffer(sim2.state_bytes(), np.float32), _state("state000999")) < 2e-6
    ru = np.fromfile(os.path.join(TT, "U00500"), np.float32).reshape(96, 96)
    assert rel(snaps2[500][0], ru) < 1e-5


def test_top_absorb_flag_is_what_the_fixture_used():
    """Without the top C-PML the step-0 psi arrays differ (the fixture came
    from the nbpsv2d_topabsorb build)."""
    n, npml = 96, 32
    vp, vs, den = O.mdl_raster(os.path.join(TT, "homo_6.8.mdl"), n, n, 0.3)
    mat, _ = O.psv_material(vp, vs, den, 0.3, 0.01)
    sim = O.PSV(n, n, npml, mat, O.psv_cpml(n, n, npml, 0.005, 1e-11, 30.0, 2.0), topabsorb=False)
    src, M = _tt_src()
    O.run_psv(sim, 1, src, M, 33, 33)
    got = np.frombuffer(sim.state_bytes(), np.float32)
    zpsi = slice(5 * n * n + 4 * 2 * npml * n, None)  # Tzz_z Txz_z U_z W_z
    assert rel(got[zpsi], _state("state000000")[zpsi]) > 0.5


@pytest.mark.slow
def test_fp64_golden_seismograms():
    """PSV/psv_clean_double/test: FP64 oracle vs the reference's FP64 run."""
    d = os.path.join(GOLD, "psv_fp64")
    nx, nz, npml, h, dt = 2048, 1024, 32, 0.3, 0.01
    vp, vs, den = O.mdl_raster(os.path.join(d, "homo.mdl"), nx, nz, h, np.float64)
    mat, stab = O.psv_material(vp, vs, den, h, dt, np.float64)
    cp = O.psv_cpml(nx, nz, npml, 0.005, 1e-11, 30.0, 2.0, np.float64)
    sim = O.PSV(nx, nz, npml, mat, cp, topabsorb=False, dtype=np.float64)
    src = O.gaussian(-5, np.float64)
    M = O.psv_moment(0, 0, 20, 90, 45, h, dt, np.float64)
    rx, rz = O.station_receivers(os.path.join(d, "station200"))
    su, sw = O.run_psv(sim, 2001, src, M, 1024, 100, rx=rx, rz=rz)
    ru = np.fromfile(os.path.join(d, "0_U"), np.float64).reshape(2001, 36)
    rw = np.fromfile(os.path.join(d, "0_W"), np.float64).reshape(2001, 36)
    assert rel(su, ru) < 1e-12 and rel(sw, rw) < 1e-12
    # the 1e-200 precursor tails agree too, trace by trace
    for i in range(36):
        if np.abs(ru[:, i]).max() > 0:
            assert rel(su[:, i], ru[:, i]) < 1e-10


def test_sh_passes_match_reference_cpu_loop():
    """One SH stress pass and one velocity pass of the oracle against the
    reference's own SH CPU loop (SH/fd/fd.cpp:45-193, compiled from the
    reference sources; vectors in tests/golden/ref_host/sh_fd.npz), from
    random fields, psi and material with the reference's C-PML profiles.
    That loop differs from the GPU kernels the oracle follows only within 4
    cells of the domain edge (order reduction without iz, no early return,
    row-major wrap), so interior cells -- including the side and bottom
    C-PML strips -- are compared, and the top imaging row in full."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_host", "sh_fd.npz"))
    nx, nz, npml = (int(v) for v in d["dims"])
    cp = {c: tuple(d["%s_%s" % (a, c)] for a in "bck") for c in ("Tx_x", "V_x", "Tz_z", "V_z")}

    def fresh():
        mat = {k: d["in_" + k] for k in ("BV", "MUX", "MUZ")}
        sim = O.SH(nx, nz, npml, mat, cp, dtype=np.float32)
        for k in ("V", "Tx", "Tz"):
            getattr(sim, k)[:] = d["in_" + k]
        for k in ("V_x", "V_z", "Tx_x", "Tz_z"):
            sim.psi[k][:] = d["in_psi_" + k]
        return sim

    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz))
    inner = (ix >= 4) & (ix <= nx - 5) & (iz >= 4) & (iz <= nz - 5)
    # psi cells: x-psi index ix + iz*2npml (left), npml+nx-1-ix + iz*2npml (right)
    xl = inner & (ix < npml)
    xr = inner & (ix >= nx - npml)
    zb = inner & (iz >= nz - npml)

    def xpsi(a):
        a2 = a.reshape(nz, 2 * npml)
        return np.concatenate([a2[:, :npml][xl[:, :npml]], a2[:, npml:][xr[:, nx - npml:][:, ::-1]]])

    def zpsi(a):
        rows = a.reshape(2 * npml, nx)
        return rows[(npml + nz - 1 - iz)[zb], ix[zb]]

    sim = fresh()
    sim.stress()
    for k in ("Tx", "Tz"):
        ref = d["stress_" + k].reshape(nz, nx)
        np.testing.assert_allclose(getattr(sim, k)[inner], ref[inner], rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(xpsi(sim.psi["V_x"]), xpsi(d["stress_psi_V_x"]), rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(zpsi(sim.psi["V_z"]), zpsi(d["stress_psi_V_z"]), rtol=2e-6, atol=2e-6)
    assert xl.any() and xr.any() and zb.any()

    sim = fresh()
    sim.velocity()
    ref = d["velocity_V"].reshape(nz, nx)
    np.testing.assert_allclose(sim.V[inner], ref[inner], rtol=2e-6, atol=2e-6)
    np.testing.assert_array_equal(sim.Tz[0], d["velocity_Tz"].reshape(nz, nx)[0])
    np.testing.assert_allclose(xpsi(sim.psi["Tx_x"]), xpsi(d["velocity_psi_Tx_x"]), rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(zpsi(sim.psi["Tz_z"]), zpsi(d["velocity_psi_Tz_z"]), rtol=2e-6, atol=2e-6)
