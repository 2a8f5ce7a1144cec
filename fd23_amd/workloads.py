"""Synthetic workloads: seeded layered + Gaussian-random-field models and the
source/receiver geometry of the benchmark configurations.

Recipe (as specified for the benchmark configurations):
  P-SV: `nlayer` flat layers with sorted uniform interface depths,
        vp ~ U(1800, 6000) m/s per layer, vs = vp/1.732,
        rho = 310 vp^0.25 (Gardner), all multiplied by (1 + sigma*GRF),
        GRF of unit variance with a Gaussian correlation length (FFT filter).
  SH:   three flat layers vs 2000/2800/3500, rho 2200/2450/2700 with
        interfaces at iz=60 and 130, vs multiplied by (1 + 0.02 g), g white.
C-PML parameters (the reference leaves them to the par file):
  pml_v = vpmax/dx, pml_dt = dt, pml_fc = f0, pml_r = 1e-6.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib, host


def grf(nz, nx, corr_cells, rng):
    """Unit-variance Gaussian random field with a Gaussian correlation of
    length corr_cells (FFT filtering of white noise)."""
    w = rng.standard_normal((nz, nx))
    kz = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    filt = np.exp(-((np.pi * corr_cells) ** 2) * (kx ** 2 + kz ** 2) / 2.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def layered_grf_model(nx, nz, dx, seed, nlayer=12, sigma=0.03, corr_m=50.0, vmin=1800.0, vmax=6000.0):
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.integers(1, nz, nlayer - 1))
    vps = rng.uniform(vmin, vmax, nlayer)
    edges = [0] + list(depths) + [nz]
    vp = np.empty((nz, nx), np.float64)
    for i in range(nlayer):
        vp[edges[i]:edges[i + 1], :] = vps[i]
    vp *= 1.0 + sigma * grf(nz, nx, corr_m / dx, rng)
    vs = vp / 1.732
    den = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), den.astype(np.float32)


def sh_three_layer_model(nx, nz, seed, sigma=0.02):
    rng = np.random.default_rng(seed)
    vs = np.empty((nz, nx), np.float64)
    den = np.empty((nz, nx), np.float64)
    for (z0, z1), v, r in (((0, 60), 2000.0, 2200.0), ((60, 130), 2800.0, 2450.0), ((130, nz), 3500.0, 2700.0)):
        vs[z0:z1] = v
        den[z0:z1] = r
    vs *= 1.0 + sigma * rng.standard_normal((nz, nx))
    return vs.astype(np.float32), den.astype(np.float32)


@dataclass
class Workload:
    name: str
    wave: int
    nx: int
    nz: int
    dx: float
    dt: float
    nt: int
    npml: int
    topabsorb: bool
    f0: float
    sources: list            # [(xs, zs)] per shot
    rec_x: np.ndarray        # receiver columns (pairs x, x-1 as the reference)
    rec_z: np.ndarray
    kind: int = _lib.SRC_LINE
    explosion: bool = True
    seed: int = 0
    storage: int = _lib.STORE_F32
    mat: list = field(default=None, repr=False)
    packed: np.ndarray = field(default=None, repr=False)
    stf: np.ndarray = field(default=None, repr=False)
    M: np.ndarray = field(default=None, repr=False)
    stab: float = 0.0
    vpmax: float = 0.0
    # first-arrival window (None: none): seconds before the straight-ray
    # first arrival at vpmax and after it
    punch_before: float = None
    punch_after: float = None

    @property
    def nshots(self):
        return len(self.sources)

    def build(self, vp=None, vs=None, den=None):
        """Material, C-PML profiles, source wavelet and moment tensor."""
        if self.wave == _lib.WAVE_PSV:
            if vp is None:
                vp, vs, den = layered_grf_model(self.nx, self.nz, self.dx, self.seed)
            f64 = self.storage == _lib.STORE_F64
            if f64:
                vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
            self.mat, self.stab = host.psv_material(vp, vs, den, self.dx, self.dt, f64=f64)
            self.vpmax = float(vp.max())
            self.M = host.psv_moment(self.kind, 0.0, 0.0, 0.0, 0.0, self.dx, self.dt, explosion=self.explosion)
        else:
            if vs is None:
                vs, den = sh_three_layer_model(self.nx, self.nz, self.seed)
            self.mat, self.stab = host.sh_material(vs, den, self.dx, self.dt)
            self.vpmax = float(vs.max())
            # vy point force with dip=90, rake=0, strike=azimuth: Mxy = 1
            self.M = host.sh_moment(_lib.SRC_LINE, 0.0, 90.0, 0.0, 0.0, self.dx, self.dt)
        self.packed = host.cpml_pack(self.wave, self.nx, self.nz, self.npml, self.dt, 1e-6,
                                     self.vpmax / self.dx, self.f0)
        self.stf = host.ricker(self.nt, self.dt, self.f0)
        return self

    def receivers(self):
        rx = np.empty(2 * self.rec_x.size, np.int32)
        rz = np.empty(2 * self.rec_x.size, np.int32)
        rx[0::2] = self.rec_x
        rx[1::2] = self.rec_x - 1
        rz[0::2] = self.rec_z
        rz[1::2] = self.rec_z
        return rx, rz

    def context(self, max_frames=None, nshots=None):
        from .core import Context

        rx, rz = self.receivers()
        S = self.nshots if nshots is None else nshots
        ctx = Context(self.wave, self.nx, self.nz, self.npml, nshots=S, storage=self.storage,
                      topabsorb=self.topabsorb, max_frames=self.nt if max_frames is None else max_frames,
                      max_rec=rx.size)
        ctx.set_material(self.mat)
        ctx.set_cpml(self.packed)
        ncomp = 2 if self.wave == _lib.WAVE_PSV else 1
        for s in range(S):
            xs, zs = self.sources[s % self.nshots]
            ctx.set_source(s, self.kind if self.wave == _lib.WAVE_PSV else _lib.SRC_FORCE, xs, zs, self.stf,
                           self.M)
            for c in range(ncomp):
                ctx.set_receivers(s, c, rx, rz)
        if self.punch_after is not None:
            ctx.set_punch(self.travel_times(), self.punch_before, self.punch_after, self.dt)
        return ctx

    def travel_times(self):
        """Straight-ray first-arrival times (s) from shot 0's source at the
        fastest speed of the model, nz x nx float32."""
        xs, zs = self.sources[0]
        z, x = np.mgrid[0:self.nz, 0:self.nx]
        return (np.hypot(x - xs, z - zs) * self.dx / self.vpmax).astype(np.float32)

    def punch_active_fraction(self, it0, nt):
        """Fraction of (32x16 block, step) pairs inside the window over
        steps [it0, it0+nt): the share of cell updates the passes perform."""
        t = self.travel_times().reshape(self.nz // 16, 16, self.nx // 32, 32)
        f32 = np.float32
        first = ((t.min(axis=(1, 3)) - f32(self.punch_before)) / f32(self.dt)).astype(np.int64)
        last = ((t.max(axis=(1, 3)) + f32(self.punch_after)) / f32(self.dt)).astype(np.int64)
        lo = np.maximum(first + 1, it0)
        hi = np.minimum(last - 1, it0 + nt - 1)
        return float(np.clip(hi - lo + 1, 0, None).sum() / (first.size * nt))

    def cell_updates(self, nt=None):
        return float(self.nshots) * self.nx * self.nz * (self.nt if nt is None else nt)


def config_sh_400x200():
    """Plain SH FP32: 400x200, 10 m, 1 ms, 2000 steps, 4-sided C-PML 20."""
    rec_x = np.arange(1, 400, 2, dtype=np.int32)
    return Workload("sh_400x200_fp32", _lib.WAVE_SH, 400, 200, 10.0, 1e-3, 2000, 20, True, 10.0,
                    [(200, 20)], rec_x, np.full(rec_x.size, 2, np.int32), kind=_lib.SRC_FORCE,
                    explosion=False, seed=1001)


def config_psv_2000x1000():
    """Plain P-SV FP32: 2000x1000, 5 m, 0.3 ms, 5000 steps, C-PML 32."""
    rec_x = np.arange(1, 2000, 2, dtype=np.int32)
    return Workload("psv_2000x1000_fp32", _lib.WAVE_PSV, 2000, 1000, 5.0, 3e-4, 5000, 32, False, 15.0,
                    [(1000, 10)], rec_x, np.full(rec_x.size, 5, np.int32), seed=2002)


def config_psv_batched(nshots=64):
    """Batched P-SV FP32: 1024x512, 5 m, 0.3 ms, 4000 steps, 64 shots."""
    xs = [32 + int(round(k * (1024 - 64 - 1) / 63)) for k in range(nshots)]
    rec_x = np.arange(1, 1024, 2, dtype=np.int32)
    return Workload("psv_1024x512_x%d_fp32" % nshots, _lib.WAVE_PSV, 1024, 512, 5.0, 3e-4, 4000, 32, False,
                    12.0, [(x, 8) for x in xs], rec_x, np.full(rec_x.size, 4, np.int32), explosion=True,
                    seed=3003)


def config_psv_precision(storage=_lib.STORE_F16, nshots=8):
    """Mixed-precision P-SV: 4096x2048, 4 m, 0.25 ms, 6000 steps, 8 shots."""
    xs = [64 + k * (4096 - 128) // 7 for k in range(nshots)]
    rec_x = np.arange(1, 4096, 2, dtype=np.int32)
    return Workload("psv_4096x2048_x%d" % nshots, _lib.WAVE_PSV, 4096, 2048, 4.0, 2.5e-4, 6000, 32, False,
                    20.0, [(x, 8) for x in xs], rec_x, np.full(rec_x.size, 4, np.int32), explosion=True,
                    seed=4004, storage=storage)


def config_psv_punch():
    """P-SV FP32 with a first-arrival window: the 2000x1000 configuration's
    physics on 2048x1024 (block multiples), 5000 steps; blocks update from
    20 ms before the straight-ray first arrival at the fastest P speed to
    300 ms after it."""
    rec_x = np.arange(1, 2048, 2, dtype=np.int32)
    return Workload("psv_2048x1024_punch_fp32", _lib.WAVE_PSV, 2048, 1024, 5.0, 3e-4, 5000, 32, False, 15.0,
                    [(1024, 10)], rec_x, np.full(rec_x.size, 5, np.int32), seed=2002, punch_before=0.02,
                    punch_after=0.3)
