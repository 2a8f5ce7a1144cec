"""Shared set-up for parity tests: one problem description, built once by
the library's host code, fed both to the GPU context and to the oracle."""
from __future__ import annotations

import numpy as np

import oracle as O
from fd23_amd import _lib, host
from fd23_amd.core import Context


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def layered_model(nx, nz, seed, vmin=1800.0, vmax=6000.0, nlayer=6, sigma=0.03):
    rng = np.random.default_rng(seed)
    depth = np.sort(rng.integers(1, nz, nlayer - 1))
    vps = rng.uniform(vmin, vmax, nlayer)
    vp = np.empty((nz, nx), np.float64)
    edges = [0] + list(depth) + [nz]
    for i in range(nlayer):
        vp[edges[i]:edges[i + 1], :] = vps[i]
    vp *= 1.0 + sigma * rng.standard_normal((nz, nx))
    vs = vp / 1.732
    den = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), den.astype(np.float32)


def unpack_cpml(packed, wave, nx, nz, dtype=np.float32):
    names = (O.PSV_CPML_X + O.PSV_CPML_Z) if wave == _lib.WAVE_PSV else ("Tx_x", "V_x", "Tz_z", "V_z")
    half = len(names) // 2
    out, off = {}, 0
    for i, name in enumerate(names):
        n = nx if i < half else nz
        out[name] = tuple(packed[off + j * n: off + (j + 1) * n].astype(dtype) for j in range(3))
        off += 3 * n
    return out


class Problem:
    """A single- or multi-shot problem with library-precomputed inputs."""

    def __init__(self, wave, nx, nz, npml, h, dt, nt, seed=1, topabsorb=False, kind=_lib.SRC_LINE,
                 explosion=False, sources=None, rec_z=5, f0=None, f64=False, vmax=6000.0):
        self.wave, self.nx, self.nz, self.npml = wave, nx, nz, npml
        self.h, self.dt, self.nt = h, dt, nt
        self.topabsorb, self.kind, self.f64 = topabsorb, kind, f64
        vp, vs, den = layered_model(nx, nz, seed, vmax=vmax)
        self.vmax = float(vp.max())
        if f64:
            vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
        if wave == _lib.WAVE_PSV:
            self.mat, self.stab = host.psv_material(vp, vs, den, h, dt, f64=f64)
        else:
            self.mat, self.stab = host.sh_material(vs, den, h, dt)
        pml_v = float(vp.max()) / h
        f0 = f0 or 12.0
        self.packed = host.cpml_pack(wave, nx, nz, npml, dt, 1e-6, pml_v, f0)
        stf = host.ricker(nt, dt, f0)
        if kind == _lib.SRC_P:
            stf = np.cumsum(stf.astype(np.float32), dtype=np.float32).astype(np.float64)
        self.stf = stf
        if wave == _lib.WAVE_PSV:
            self.M = host.psv_moment(kind, 30.0, 60.0, 45.0, 10.0, h, dt, explosion=explosion)
        else:
            self.M = host.sh_moment(kind if kind != _lib.SRC_FORCE else _lib.SRC_LINE, 30.0, 60.0, 45.0,
                                    10.0, h, dt)
        self.sources = sources or [(nx // 2, 10)]
        self.rx = np.arange(1, nx, 2, dtype=np.int32)
        self.rz = np.full(self.rx.size, rec_z, np.int32)

    @property
    def nshots(self):
        return len(self.sources)

    def context(self, storage=None, max_frames=None):
        if storage is None:
            storage = _lib.STORE_F64 if self.f64 else _lib.STORE_F32
        ctx = Context(self.wave, self.nx, self.nz, self.npml, nshots=self.nshots, storage=storage,
                      topabsorb=self.topabsorb, max_frames=self.nt if max_frames is None else max_frames,
                      max_rec=self.rx.size)
        ctx.set_material(self.mat)
        ctx.set_cpml(self.packed)
        ncomp = 2 if self.wave == _lib.WAVE_PSV else 1
        for s, (xs, zs) in enumerate(self.sources):
            ctx.set_source(s, self.kind, xs, zs, self.stf, self.M)
            for c in range(ncomp):
                ctx.set_receivers(s, c, self.rx, self.rz)
        return ctx

    def travel_times(self, shot=0, v=None):
        """Straight-ray first-arrival times from a shot's source at speed v
        (default: the fastest P speed of the model), nz x nx float32."""
        xs, zs = self.sources[shot]
        z, x = np.mgrid[0:self.nz, 0:self.nx]
        return (np.hypot(x - xs, z - zs) * self.h / (v or self.vmax)).astype(np.float32)

    def oracle(self, shot=0, window=None):
        """window: (times, before, after) for a first-arrival window."""
        dt_ = np.float64 if self.f64 else np.float32
        cp = unpack_cpml(self.packed, self.wave, self.nx, self.nz, dt_)
        xs, zs = self.sources[shot]
        if self.wave == _lib.WAVE_PSV:
            mat = dict(zip(("BU", "BW", "MU", "MUA", "LAM"), self.mat))
            sim = O.PSV(self.nx, self.nz, self.npml, mat, cp, self.topabsorb, dt_)
            if window is not None:
                sim.set_window(*window, self.dt)
            okind = {(_lib.SRC_LINE, False): 0, (_lib.SRC_P, False): 1}.get((self.kind, False), 0)
            u, w = O.run_psv(sim, self.nt, self.stf.astype(dt_), self.M.astype(dt_), xs, zs, kind=okind,
                             rx=self.rx, rz=self.rz)
            return sim, [u, w]
        mat = dict(zip(("BV", "MUX", "MUZ"), self.mat))
        sim = O.SH(self.nx, self.nz, self.npml, mat, cp, self.topabsorb, dt_)
        if window is not None:
            sim.set_window(*window, self.dt)
        v = O.run_sh(sim, self.nt, self.stf.astype(dt_), self.M.astype(dt_), xs, zs, kind=self.kind,
                     rx=self.rx, rz=self.rz)
        return sim, [v]


# --------------------------------------------------------------------------
# reduced-precision storage emulated on the FP64 oracle: after every pass the
# arrays that pass updated are rounded to the storage format (round to
# nearest even, one power-of-two scale per array).  This is the error floor
# of the format itself; the GPU's FP16/BF16 path is checked against it.
# --------------------------------------------------------------------------
def round_bf16(a):
    """float64 -> nearest bfloat16 (ties to even) -> float64; finite input."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def round_storage(a, fmt):
    """Round array a in place as stored in fmt ("f16", "bf16", "f32") with a
    power-of-two scale that puts its largest magnitude in [2^8, 2^9)."""
    m = float(np.abs(a).max())
    if m == 0.0:
        return
    if fmt == "f32":
        a[...] = a.astype(np.float32)
        return
    s = 2.0 ** (int(np.floor(np.log2(m))) - 8)
    if fmt == "f16":
        a[...] = (a / s).astype(np.float16).astype(np.float64) * s
    elif fmt == "bf16":
        a[...] = round_bf16(a / s) * s
    else:
        raise ValueError(fmt)


def emulated_storage(p: Problem, fmt):
    """Seismograms of p's shot 0 with FP64 arithmetic and fmt storage (every
    array a pass updates rounded after the pass; the oracle's C rounding,
    bitwise equal to round_storage)."""
    cp = unpack_cpml(p.packed, p.wave, p.nx, p.nz, np.float64)
    xs, zs = p.sources[0]
    stf, M = p.stf.astype(np.float64), p.M.astype(np.float64)
    if p.wave == _lib.WAVE_PSV:
        sim = O.PSV(p.nx, p.nz, p.npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), p.mat)), cp, p.topabsorb,
                    np.float64)
        after_stress = lambda: [sim.Txx, sim.Tzz, sim.Txz] + [sim.psi[n] for n in ("U_x", "W_x", "U_z", "W_z")]
        after_vel = lambda: [sim.U, sim.W] + [sim.psi[n] for n in ("Txx_x", "Txz_x", "Txz_z", "Tzz_z")]
        comps, kind = ("U", "W"), {_lib.SRC_LINE: 0, _lib.SRC_P: 1}[p.kind]
    else:
        sim = O.SH(p.nx, p.nz, p.npml, dict(zip(("BV", "MUX", "MUZ"), p.mat)), cp, p.topabsorb, np.float64)
        after_stress = lambda: [sim.Tx, sim.Tz, sim.psi["V_x"], sim.psi["V_z"]]
        after_vel = lambda: [sim.V, sim.psi["Tx_x"], sim.psi["Tz_z"]]
        comps, kind = ("V",), p.kind
    seis = [[] for _ in comps]
    for it in range(p.nt):
        sim.stress()
        sim.inject(kind, it, stf, M, xs, zs)
        for a in after_stress():
            O.round_storage(a, fmt)
        sim.velocity()
        for a in after_vel():
            O.round_storage(a, fmt)
        for c, name in enumerate(comps):
            seis[c].append(sim.record(name, p.rx, p.rz))
    return [np.array(s) for s in seis]


def point_source_frames(sim, nt, xs, zs, bw, stf, M):
    """Displacement frames for the file-driven box (the sourcefile layout of
    BOX::setboxsource: per frame U then W on the (bw+8)^2 window around the
    bw x bw box at (xs, zs)) from a run of sim with a point source at
    (xs, zs) and no box: frame k = dt * (sum of U, W after steps 0..k-2),
    frames 0..nt.  Step it's drive is then the velocity before step it and
    C grad(frame it+1)/dt the stress after its stress pass, wherever no
    source acts.  sim is stepped nt times (its fields end as the run's)."""
    s0 = (bw + 1) // 2 + 1 - bw
    wx = slice(xs + s0 - 4, xs + s0 + bw + 4)
    wz = slice(zs + s0 - 4, zs + s0 + bw + 4)
    dt = sim.dt_frames
    frames = np.zeros((nt + 1, 2, bw + 8, bw + 8))
    acc = np.zeros((2, bw + 8, bw + 8))
    for it in range(nt):
        sim.stress(it)
        sim.inject(0, it, stf, M, xs, zs)
        sim.velocity(it)
        if it + 2 <= nt:
            acc = acc + dt * np.stack([sim.U[wz, wx], sim.W[wz, wx]]).astype(np.float64)
            frames[it + 2] = acc
    return frames
