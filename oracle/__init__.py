"""Oracle for the fd23 staggered-grid elastic loop -- TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference's time-stepping loop with the semantics
of its GPU kernels, in FP32 and FP64 (C in ``fd_oracle.c``, driven from
here).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker.

Reference files followed (paths relative to the reference root):
  stress / velocity passes   PSV/clean_nobox/fd/cu_fd.cu:96-388, SH/fd/cu_fd.cu:77-273
  C-PML profiles             PSV/clean_nobox/cpml/cpml.cpp:7-186, SH/cpml/cpml.cpp:7-177
  material                   PSV/clean_nobox/material/material.cpp:35-240,
                             SH/material/material.cpp:31-135
  point sources              PSV/clean_nobox/point_source/cu_point_source.cu:6-159,
                             SH/point_source/cu_point_source.cu:6-201
  source-time functions      PSV/clean_nobox/source/waveform.cpp:8-47
  receivers                  SH/record/cu_record.cu:12-77,136-144
  loop order                 PSV/clean_nobox/main.cpp:115-168, SH/main.cpp:93-141
  state%06d layout           PSV/clean_nobox/main.cpp:253-318
  .mdl models                SH/model/model.cpp:175-469
  first-arrival window       SH/slide/cu_slide.cu:7-61, SH/fd/cu_fd.cu:18-25,
                             PSV/psv_clean_double/fd/cu_fd.cu:16-23, main.cpp:120
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PSV_CPML_X = ("Txx_x", "Txz_x", "U_x", "W_x")
PSV_CPML_Z = ("Tzz_z", "Txz_z", "U_z", "W_z")
# staggering offsets, PSV/clean_nobox/cpml/cpml.cpp:177-185
PSV_CPML_START = {"W_x": 0.5, "Txx_x": 0.5, "U_x": 0.0, "Txz_x": 0.0,
                  "W_z": 0.5, "Txz_z": 0.5, "U_z": 0.0, "Tzz_z": 0.0}
# SH/cpml/cpml.cpp:172-176
SH_CPML_START = {"Tx_x": 0.0, "Tz_z": 0.5, "V_x": 0.5, "V_z": 0.0}


def build(force: bool = False) -> str:
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = ctypes.CDLL(so)
    return _LIB


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ptrs(arrs):
    t = ctypes.c_void_p * len(arrs)
    return t(*[a.ctypes.data for a in arrs])


# --------------------------------------------------------------------------
# host precompute
# --------------------------------------------------------------------------
def cpml_profile(n, npml, startpos, pml_dt, pml_r, pml_v, pml_fc, dtype=np.float32):
    """b, c, k(=1/kappa) of one 1-D C-PML profile (cpml.cpp:7-83)."""
    b = np.zeros(n, dtype)
    c = np.zeros(n, dtype)
    k = np.zeros(n, dtype)
    f = getattr(lib(), "or_cpml_coef" + _sfx(dtype))
    f.restype = None
    f(ctypes.c_int(n), ctypes.c_int(npml), ctypes.c_double(npml - 0.5),
      ctypes.c_double(startpos), ctypes.c_double(pml_dt), ctypes.c_double(pml_r),
      ctypes.c_double(pml_v), ctypes.c_double(pml_fc), _p(b), _p(c), _p(k))
    return b, c, k


def psv_cpml(nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc, dtype=np.float32):
    out = {}
    for name, st in PSV_CPML_START.items():
        n = nx if name.endswith("_x") else nz
        out[name] = cpml_profile(n, npml, st, pml_dt, pml_r, pml_v, pml_fc, dtype)
    return out


def sh_cpml(nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc, dtype=np.float32):
    out = {}
    for name, st in SH_CPML_START.items():
        n = nx if name.endswith("_x") else nz
        out[name] = cpml_profile(n, npml, st, pml_dt, pml_r, pml_v, pml_fc, dtype)
    return out


def psv_material(vp, vs, den, h, dt, dtype=np.float32):
    """Returns (dict BU,BW,MU,MUA,LAM, stab)."""
    nz, nx = vp.shape
    vp = np.ascontiguousarray(vp, dtype)
    vs = np.ascontiguousarray(vs, dtype)
    den = np.ascontiguousarray(den, dtype)
    out = {k: np.zeros((nz, nx), dtype) for k in ("BU", "BW", "MU", "MUA", "LAM")}
    f = getattr(lib(), "or_psv_material" + _sfx(dtype))
    f.restype = ctypes.c_double
    stab = f(ctypes.c_int(nx), ctypes.c_int(nz), ctypes.c_double(h), ctypes.c_double(dt),
             _p(vp), _p(vs), _p(den), _p(out["BU"]), _p(out["BW"]), _p(out["MU"]),
             _p(out["MUA"]), _p(out["LAM"]))
    return out, stab


def sh_material(vs, den, h, dt, dtype=np.float32):
    nz, nx = vs.shape
    vs = np.ascontiguousarray(vs, dtype)
    den = np.ascontiguousarray(den, dtype)
    out = {k: np.zeros((nz, nx), dtype) for k in ("BV", "MUX", "MUZ")}
    f = getattr(lib(), "or_sh_material" + _sfx(dtype))
    f.restype = ctypes.c_double
    stab = f(ctypes.c_int(nx), ctypes.c_int(nz), ctypes.c_double(h), ctypes.c_double(dt),
             _p(vs), _p(den), _p(out["BV"]), _p(out["MUX"]), _p(out["MUZ"]))
    return out, stab


def gaussian(alpha, dtype=np.float32):
    """waveform.cpp:27-47; alpha is truncated to int first (Q14)."""
    a = int(alpha)
    n = 2 * abs(int(a * 3.0)) + 2
    out = np.zeros(n, dtype)
    f = getattr(lib(), "or_gaussian" + _sfx(dtype))
    f.restype = ctypes.c_int
    lsrc = f(ctypes.c_int(a), _p(out))
    return out[:lsrc].copy()


def trapezoid(dt, trap1, trap2, trap3, dtype=np.float32):
    n = int(trap1 / dt) + int(trap2 / dt) + int(trap3 / dt) + 8
    out = np.zeros(n, dtype)
    f = getattr(lib(), "or_trapezoid" + _sfx(dtype))
    f.restype = ctypes.c_int
    lsrc = f(ctypes.c_double(dt), ctypes.c_double(trap1), ctypes.c_double(trap2),
             ctypes.c_double(trap3), _p(out))
    return out[:lsrc].copy()


def ricker(n, dt, f0, t0=None):
    if t0 is None:
        t0 = 1.2 / f0
    out = np.zeros(n, np.float64)
    f = lib().or_ricker
    f.restype = None
    f(ctypes.c_int(n), ctypes.c_double(dt), ctypes.c_double(f0), ctypes.c_double(t0), _p(out))
    return out


def round_storage(a, fmt):
    """Round the float64 array a in place as stored in fmt ("f32", "f16",
    "bf16") with one power-of-two scale putting its largest magnitude in
    [2^8, 2^9), or as block-scaled int16 ("blk16": one power-of-two
    quantum per 64 consecutive values) (or_round_storage; the storage
    floor's emulation)."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    f = lib().or_round_storage
    f.restype = None
    f(_p(a), ctypes.c_long(a.size), ctypes.c_int({"f32": 0, "f16": 1, "bf16": 2, "blk16": 3}[fmt]))


def psv_moment(kind, strike, dip, rake, azimuth, h, dt, dtype=np.float32):
    M = np.zeros(3, dtype)
    f = getattr(lib(), "or_psv_moment" + _sfx(dtype))
    f.restype = None
    f(ctypes.c_int(kind), ctypes.c_double(strike), ctypes.c_double(dip),
      ctypes.c_double(rake), ctypes.c_double(azimuth), ctypes.c_double(h),
      ctypes.c_double(dt), _p(M))
    return M


def sh_moment(kind, strike, dip, rake, azimuth, h, dt, dtype=np.float32):
    M = np.zeros(2, dtype)
    f = getattr(lib(), "or_sh_moment" + _sfx(dtype))
    f.restype = None
    f(ctypes.c_int(kind), ctypes.c_double(strike), ctypes.c_double(dip),
      ctypes.c_double(rake), ctypes.c_double(azimuth), ctypes.c_double(h),
      ctypes.c_double(dt), _p(M))
    return M


def line_receivers(nrec, ix0, iz0, idx, idz):
    """RECORD::init_line, SH/record/cu_record.cu:12-28: pairs at x and x-1."""
    rx = np.zeros(2 * nrec, np.int32)
    rz = np.zeros(2 * nrec, np.int32)
    for i in range(nrec):
        rx[2 * i] = ix0 + idx * i
        rz[2 * i] = iz0 + idz * i
        rx[2 * i + 1] = rx[2 * i] - 1
        rz[2 * i + 1] = rz[2 * i]
    return rx, rz


def station_receivers(path):
    """RECORD::init_list, SH/record/cu_record.cu:30-77: 'z x' pairs."""
    vals = [int(t) for t in open(path).read().split()]
    zs, xs = vals[0::2], vals[1::2]
    rx = np.zeros(2 * len(xs), np.int32)
    rz = np.zeros(2 * len(xs), np.int32)
    rx[0::2] = xs
    rz[0::2] = zs
    rx[1::2] = np.asarray(xs) - 1
    rz[1::2] = zs
    return rx, rz


# --------------------------------------------------------------------------
# .mdl polygon models (SH/model/model.cpp:175-469), vectorised
# --------------------------------------------------------------------------
def load_mdl(path):
    polys = []
    defs = {}
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if not line or line.startswith("#"):
            continue
        args = line.split()
        if not args:
            continue
        key = args[0].upper()
        if key == "DEF":
            defs[args[1]] = (float(args[2]), float(args[3]))
        elif key in ("BACK", "POLY"):
            vp, vs, den = (float(a) for a in args[1:4])
            if vs < 0:
                vs = -vp * vs
            if den < 0:
                den = -vp * den
            verts = []
            if key == "POLY":
                while i < len(lines) and lines[i].startswith("\t"):
                    toks = lines[i].split()
                    i += 1
                    j = 0
                    while j < len(toks):
                        if toks[j][0].isalpha():
                            verts.append(defs[toks[j]])
                            j += 1
                        else:
                            verts.append((float(toks[j]), float(toks[j + 1])))
                            j += 2
                polys.append((vp, vs, den, verts))
            else:
                polys.insert(0, (vp, vs, den, None))
        else:
            raise ValueError("bad model command: %r" % line)
    return polys


def _inside(xp, zp, verts, vdtype=np.float32):
    """Crossing-number test of model.cpp:437-469, over arrays of points."""
    v = np.asarray(verts, vdtype).astype(np.float64)
    n = len(v)
    res = np.zeros(xp.shape, np.int64)
    hit = np.zeros(xp.shape, bool)
    x2, y2 = v[n - 1]
    for j in range(n):
        x1, y1 = x2, y2
        x2, y2 = v[j]
        act = ~((min(y1, y2) >= zp) | (max(y1, y2) < zp)) & ~hit
        if y1 == y2:
            gt = xp > max(x1, x2)
            on = (~gt) & (xp >= min(x1, x2))
            hit |= act & on
            res += (act & ~gt & ~on).astype(np.int64)
        else:
            xi = x1 + (zp - y1) * (x2 - x1) / (y2 - y1)
            hit |= act & (xi == xp)
            res += (act & (xi != xp) & (xi > xp)).astype(np.int64)
    return hit | (res % 2 == 1)


def mdl_raster(path, nx, nz, h, dtype=np.float32):
    """get_elas_model: background then polygons in order, later ones win.

    The FP32 trees keep vertices and velocities as float
    (SH/model/model.cpp:12-26); the FP64 tree keeps them as double
    (PSV/psv_clean_double/model/model.cpp:12-26): dtype selects which."""
    polys = load_mdl(path)
    pd = np.float64 if np.dtype(dtype) == np.float64 else np.float32
    vp = np.full((nz, nx), pd(polys[0][0]), np.float64)
    vs = np.full((nz, nx), pd(polys[0][1]), np.float64)
    den = np.full((nz, nx), pd(polys[0][2]), np.float64)
    hh = float(np.float32(h)) if np.dtype(dtype) == np.float32 else float(h)
    for pvp, pvs, pden, verts in polys[1:]:
        va = np.asarray(verts, np.float32)
        x1, x2 = va[:, 0].min(), va[:, 0].max()
        z1, z2 = va[:, 1].min(), va[:, 1].max()
        if pd == np.float64:
            va = np.asarray(verts, np.float64)
            x1, x2 = np.float32(va[:, 0].min()), np.float32(va[:, 0].max())
            z1, z2 = np.float32(va[:, 1].min()), np.float32(va[:, 1].max())
        ix1 = min(max(int(x1 / hh - 1), 0), nx - 1)
        ix2 = min(max(int(x2 / hh + 1), 0), nx - 1)
        iz1 = min(max(int(z1 / hh - 1), 0), nz - 1)
        iz2 = min(max(int(z2 / hh + 1), 0), nz - 1)
        zz, xx = np.meshgrid(np.arange(iz1, iz2 + 1), np.arange(ix1, ix2 + 1), indexing="ij")
        m = _inside(hh * xx, hh * zz, verts, pd)
        sl = (slice(iz1, iz2 + 1), slice(ix1, ix2 + 1))
        vp[sl][m] = pd(pvp)
        vs[sl][m] = pd(pvs)
        den[sl][m] = pd(pden)
    return vp.astype(dtype), vs.astype(dtype), den.astype(dtype)


# --------------------------------------------------------------------------
# first-arrival window ("punch")
# --------------------------------------------------------------------------
PUNCH_BX, PUNCH_BZ = 32, 16


def punch_window(times, before, after, dt, dtype=np.float32):
    """Per 32 x 16 block, the first and last step of the window,
    SH/slide/cu_slide.cu:27-53: the block's min and max of the travel-time
    grid (nz x nx), (min - before) / dt and (max + after) / dt, truncated
    toward zero.  dtype float32 follows the FP32 trees; float64 the FP64
    tree, whose timefile, before, after and dt are double
    (PSV/psv_clean_double/slide/cu_slide.cu:10-44)."""
    t = np.ascontiguousarray(times, dtype)
    nz, nx = t.shape
    nzb, nxb = nz // PUNCH_BZ, nx // PUNCH_BX
    blk = t[:nzb * PUNCH_BZ, :nxb * PUNCH_BX].reshape(nzb, PUNCH_BZ, nxb, PUNCH_BX)
    lo = blk.min(axis=(1, 3))
    hi = blk.max(axis=(1, 3))
    f = np.dtype(dtype).type
    first = ((lo - f(before)) / f(dt)).astype(np.int32)
    last = ((hi + f(after)) / f(dt)).astype(np.int32)
    return first, last


def punch_mask(first, last, it):
    """Per cell, whether the passes update it at step it: first < it < last
    of its block (the kernels return otherwise, SH/fd/cu_fd.cu:18-25)."""
    open_b = (first < it) & (it < last)
    return np.ascontiguousarray(np.repeat(np.repeat(open_b, PUNCH_BZ, axis=0), PUNCH_BX, axis=1)
                                .astype(np.uint8))


# --------------------------------------------------------------------------
# stepping
# --------------------------------------------------------------------------
class _Window:
    """First-arrival window shared by the P-SV and SH state classes: with a
    window set, each pass takes the step index it."""
    window = None

    def set_window(self, times, before, after, dt):
        if times is None:
            self.window = None
            return
        if self.nx % PUNCH_BX or self.nz % PUNCH_BZ:
            raise ValueError("first-arrival window needs nx % 32 == 0 and nz % 16 == 0")
        if np.shape(times) != (self.nz, self.nx):
            raise ValueError(f"travel times {np.shape(times)} != (nz, nx)")
        self.window = punch_window(times, before, after, dt)

    def _open(self, it):
        if self.window is None:
            return None
        if it is None:
            raise ValueError("a first-arrival window is set: pass the step index")
        m = punch_mask(*self.window, it)
        self._mask = m
        return m.ctypes.data_as(ctypes.c_void_p)


class PSV(_Window):
    """State and passes of the P-SV loop at one precision."""

    def __init__(self, nx, nz, npml, mat, cpml, topabsorb=False, dtype=np.float32):
        self.nx, self.nz, self.npml = nx, nz, npml
        self.topabsorb = int(bool(topabsorb))
        self.dtype = np.dtype(dtype)
        z = lambda *s: np.zeros(s, dtype)
        self.U, self.W, self.Txx, self.Txz, self.Tzz = (z(nz, nx) for _ in range(5))
        self.psi = {n: z(2 * npml * nz) for n in PSV_CPML_X}
        self.psi.update({n: z(2 * npml * nx) for n in PSV_CPML_Z})
        self.mat = {k: np.ascontiguousarray(v, dtype) for k, v in mat.items()}
        self.cpml = {k: tuple(np.ascontiguousarray(a, dtype) for a in v) for k, v in cpml.items()}
        s = _sfx(dtype)
        L = lib()
        self._stress = getattr(L, "or_psv_stress" + s)
        self._vel = getattr(L, "or_psv_velocity" + s)
        self._inj = getattr(L, "or_psv_inject" + s)
        self._rec = getattr(L, "or_record" + s)
        for f in (self._stress, self._vel, self._inj, self._rec):
            f.restype = None

    def stress(self, it=None):
        names = ("U_x", "W_x", "U_z", "W_z")
        ps = [self.psi[n] for n in names]
        bs = [self.cpml[n][0] for n in names]
        cs = [self.cpml[n][1] for n in names]
        ks = [self.cpml[n][2] for n in names]
        m = self.mat
        self._stress(ctypes.c_int(self.nx), ctypes.c_int(self.nz), ctypes.c_int(self.npml),
                     ctypes.c_int(self.topabsorb), _p(self.U), _p(self.W), _p(self.Txx),
                     _p(self.Tzz), _p(self.Txz), _p(m["LAM"]), _p(m["MU"]), _p(m["MUA"]),
                     _ptrs(ps), _ptrs(bs), _ptrs(cs), _ptrs(ks), self._open(it))

    def velocity(self, it=None):
        names = ("Txx_x", "Txz_x", "Txz_z", "Tzz_z")
        ps = [self.psi[n] for n in names]
        bs = [self.cpml[n][0] for n in names]
        cs = [self.cpml[n][1] for n in names]
        ks = [self.cpml[n][2] for n in names]
        m = self.mat
        self._vel(ctypes.c_int(self.nx), ctypes.c_int(self.nz), ctypes.c_int(self.npml),
                  ctypes.c_int(self.topabsorb), _p(self.Txx), _p(self.Tzz), _p(self.Txz),
                  _p(self.U), _p(self.W), _p(m["BU"]), _p(m["BW"]),
                  _ptrs(ps), _ptrs(bs), _ptrs(cs), _ptrs(ks), self._open(it))

    def inject(self, kind, it, src, M, xs, zs):
        src = np.ascontiguousarray(src, self.dtype)
        M = np.ascontiguousarray(M, self.dtype)
        self._inj(ctypes.c_int(kind), ctypes.c_int(it), ctypes.c_int(len(src)), _p(src), _p(M),
                  ctypes.c_int(xs), ctypes.c_int(zs), ctypes.c_int(self.nx), ctypes.c_int(self.nz),
                  _p(self.Txx), _p(self.Tzz), _p(self.Txz))

    def record(self, field, rx, rz):
        out = np.zeros(len(rx), self.dtype)
        f = getattr(self, field)
        self._rec(ctypes.c_int(self.nx), ctypes.c_int(self.nz), _p(f), ctypes.c_int(len(rx)),
                  _p(np.ascontiguousarray(rx, np.int32)), _p(np.ascontiguousarray(rz, np.int32)),
                  _p(out))
        return out

    # state%06d layout, PSV/clean_nobox/main.cpp:299-316
    def state_bytes(self):
        parts = [self.U, self.W, self.Txx, self.Txz, self.Tzz]
        parts += [self.psi[n] for n in PSV_CPML_X] + [self.psi[n] for n in PSV_CPML_Z]
        return b"".join(np.ascontiguousarray(p).tobytes() for p in parts)

    def load_state(self, buf):
        a = np.frombuffer(buf, dtype=self.dtype)
        N = self.nx * self.nz
        off = 0
        for name in ("U", "W", "Txx", "Txz", "Tzz"):
            getattr(self, name)[...] = a[off:off + N].reshape(self.nz, self.nx)
            off += N
        for name in PSV_CPML_X:
            n = 2 * self.npml * self.nz
            self.psi[name][...] = a[off:off + n]
            off += n
        for name in PSV_CPML_Z:
            n = 2 * self.npml * self.nx
            self.psi[name][...] = a[off:off + n]
            off += n


class SH(_Window):
    def __init__(self, nx, nz, npml, mat, cpml, topabsorb=False, dtype=np.float32):
        self.nx, self.nz, self.npml = nx, nz, npml
        self.topabsorb = int(bool(topabsorb))
        self.dtype = np.dtype(dtype)
        z = lambda *s: np.zeros(s, dtype)
        self.V, self.Tx, self.Tz = (z(nz, nx) for _ in range(3))
        self.psi = {"Tx_x": z(2 * npml * nz), "V_x": z(2 * npml * nz),
                    "Tz_z": z(2 * npml * nx), "V_z": z(2 * npml * nx)}
        self.mat = {k: np.ascontiguousarray(v, dtype) for k, v in mat.items()}
        self.cpml = {k: tuple(np.ascontiguousarray(a, dtype) for a in v) for k, v in cpml.items()}
        s = _sfx(dtype)
        L = lib()
        self._stress = getattr(L, "or_sh_stress" + s)
        self._vel = getattr(L, "or_sh_velocity" + s)
        self._inj = getattr(L, "or_sh_inject" + s)
        self._rec = getattr(L, "or_record" + s)
        for f in (self._stress, self._vel, self._inj, self._rec):
            f.restype = None

    def stress(self, it=None):
        names = ("V_x", "V_z")
        m = self.mat
        self._stress(ctypes.c_int(self.nx), ctypes.c_int(self.nz), ctypes.c_int(self.npml),
                     ctypes.c_int(self.topabsorb), _p(self.V), _p(self.Tx), _p(self.Tz),
                     _p(m["MUX"]), _p(m["MUZ"]), _ptrs([self.psi[n] for n in names]),
                     _ptrs([self.cpml[n][0] for n in names]),
                     _ptrs([self.cpml[n][1] for n in names]),
                     _ptrs([self.cpml[n][2] for n in names]), self._open(it))

    def velocity(self, it=None):
        names = ("Tx_x", "Tz_z")
        m = self.mat
        self._vel(ctypes.c_int(self.nx), ctypes.c_int(self.nz), ctypes.c_int(self.npml),
                  ctypes.c_int(self.topabsorb), _p(self.Tx), _p(self.Tz), _p(self.V),
                  _p(m["BV"]), _ptrs([self.psi[n] for n in names]),
                  _ptrs([self.cpml[n][0] for n in names]),
                  _ptrs([self.cpml[n][1] for n in names]),
                  _ptrs([self.cpml[n][2] for n in names]), self._open(it))

    def inject(self, kind, it, src, M, xs, zs):
        src = np.ascontiguousarray(src, self.dtype)
        M = np.ascontiguousarray(M, self.dtype)
        self._inj(ctypes.c_int(kind), ctypes.c_int(it), ctypes.c_int(len(src)), _p(src), _p(M),
                  ctypes.c_int(xs), ctypes.c_int(zs), ctypes.c_int(self.nx), ctypes.c_int(self.nz),
                  _p(self.Tx), _p(self.Tz), _p(self.V))

    def record(self, field, rx, rz):
        out = np.zeros(len(rx), self.dtype)
        f = getattr(self, field)
        self._rec(ctypes.c_int(self.nx), ctypes.c_int(self.nz), _p(f), ctypes.c_int(len(rx)),
                  _p(np.ascontiguousarray(rx, np.int32)), _p(np.ascontiguousarray(rz, np.int32)),
                  _p(out))
        return out


def run_psv(sim, nt, src, M, xs, zs, kind=0, rx=None, rz=None, rxw=None, rzw=None,
            it0=0, itrecord=1, snap_cb=None):
    """The per-step order of PSV/clean_nobox/main.cpp:115-168.

    src: the STF (kind 0/2) or its running sum (kind 1/3).  Returns the
    time-major seismograms (U, W) as [nrecorded, nrec] arrays."""
    su, sw = [], []
    if rxw is None:
        rxw, rzw = rx, rz
    for it in range(it0, it0 + nt):
        sim.stress(it)
        sim.inject(kind, it, src, M, xs, zs)
        sim.velocity(it)
        if rx is not None and it % itrecord == 0:
            su.append(sim.record("U", rx, rz))
            sw.append(sim.record("W", rxw, rzw))
        if snap_cb is not None:
            snap_cb(it, sim)
    if rx is None:
        return None, None
    return np.array(su), np.array(sw)


def run_sh(sim, nt, src, M, xs, zs, kind=0, rx=None, rz=None, itrecord=1, it0=0):
    """The per-step order of SH/main.cpp:93-141; returns V seismograms."""
    sv = []
    for it in range(it0, it0 + nt):
        sim.stress(it)
        sim.inject(kind, it, src, M, xs, zs)
        sim.velocity(it)
        if rx is not None and it % itrecord == 0:
            sv.append(sim.record("V", rx, rz))
    return np.array(sv) if rx is not None else None


# --------------------------------------------------------------------------
# hybrid box coupling (PSV/psv_clean_double/box/box.cpp:10-210 ring
# arithmetic, fd/cu_fd.cu:93-207 residual grid, main.cpp:153-200 order),
# with a rectangular box and an analytic plane-wave drive s = -incident
# --------------------------------------------------------------------------
class PlaneWaveBox:
    """Box cells [x0, x0+bx) x [z0, z0+bz); incident 0 = P, 1 = SV."""

    def __init__(self, sim, x0, z0, bx, bz, incident, angle_deg, f0, t0, amp, vp, vs, rho, h, dt):
        self.sim, self.x0, self.z0, self.bx, self.bz = sim, x0, z0, bx, bz
        self.h, self.dt, self.f0, self.t0, self.amp = h, dt, f0, t0, amp
        th = np.deg2rad(angle_deg)
        self.n = np.array([np.sin(th), -np.cos(th)])
        mu = rho * vs * vs
        lam = rho * vp * vp - 2.0 * mu
        if incident == 0:
            self.q, self.c = self.n.copy(), vp
            nx_, nz_ = self.n
            self.sfac = {"Txx": -(lam + 2 * mu * nx_ * nx_) / vp, "Tzz": -(lam + 2 * mu * nz_ * nz_) / vp,
                         "Txz": -(2 * mu * nx_ * nz_) / vp}
        else:
            self.q, self.c = np.array([np.cos(th), np.sin(th)]), vs
            (qx, qz), (nx_, nz_) = self.q, self.n
            self.sfac = {"Txx": -2 * mu * qx * nx_ / vs, "Tzz": -2 * mu * qz * nz_ / vs,
                         "Txz": -mu * (qx * nz_ + qz * nx_) / vs}
        corners = [self.n[0] * (x0 + cx * (bx - 1)) * h + self.n[1] * (z0 + cz * (bz - 1)) * h
                   for cx in (0, 1) for cz in (0, 1)]
        self.nref = min(corners) - 1.0 * h
        dt_ = sim.dtype
        fac = dt / h
        self.lam, self.mu, self.b = (dt_.type(lam * fac), dt_.type(mu * fac), dt_.type(fac / rho))
        sl = (slice(z0, z0 + bz), slice(x0, x0 + bx))
        for k, v in (("BU", self.b), ("BW", self.b), ("MU", self.mu), ("MUA", self.mu), ("LAM", self.lam)):
            sim.mat[k][sl] = v
        self.r = {k: np.zeros((bz, bx), dt_) for k in ("U", "W", "Txx", "Tzz", "Txz")}
        j, i = np.mgrid[0:bz, 0:bx]
        d = np.minimum(np.minimum(i, j), np.minimum(bx - 1 - i, bz - 1 - j))
        self.sub, self.add = (d <= 3), (d >= 4) & (d <= 7)
        self.i, self.j = i, j
        self.rstress = getattr(lib(), "or_box_stress" + _sfx(dt_))
        self.rvel = getattr(lib(), "or_box_velocity" + _sfx(dt_))
        self.rstress.restype = self.rvel.restype = None

    OFF = {"U": (0.5, 0.0), "W": (0.0, -0.5), "Txx": (0.0, 0.0), "Tzz": (0.0, 0.0), "Txz": (0.5, -0.5)}

    def drive(self, name, t):
        ox, oz = self.OFF[name]
        x = (self.x0 + self.i + ox) * self.h
        z = (self.z0 + self.j + oz) * self.h
        tau = t - (self.n[0] * x + self.n[1] * z - self.nref) / self.c - self.t0
        a = (np.pi * self.f0 * tau) ** 2
        g = -self.amp * (1.0 - 2.0 * a) * np.exp(-a)
        if name == "U":
            v = g * self.q[0]
        elif name == "W":
            v = g * self.q[1]
        else:
            v = g * self.sfac[name]
        return v.astype(self.sim.dtype)

    def _ring(self, names, t, add):
        sl = (slice(self.z0, self.z0 + self.bz), slice(self.x0, self.x0 + self.bx))
        m = self.add if add else self.sub
        for nm in names:
            s = self.drive(nm, t)
            a = getattr(self.sim, nm)[sl]
            if add:
                a[m] = self.r[nm][m] + s[m]
            else:
                self.r[nm][m] = a[m] - s[m]

    def pre_stress(self, it):
        self._ring(("U", "W"), (it + 0.5) * self.dt, True)

    def post_stress(self, it):
        self._ring(("U", "W"), (it + 0.5) * self.dt, False)
        r = self.r
        self.rstress(ctypes.c_int(self.bx), ctypes.c_int(self.bz), _p(r["U"]), _p(r["W"]), _p(r["Txx"]),
                     _p(r["Tzz"]), _p(r["Txz"]), _p(np.full((self.bz, self.bx), self.lam)),
                     _p(np.full((self.bz, self.bx), self.mu)), _p(np.full((self.bz, self.bx), self.mu)))
        self._ring(("Txx", "Tzz", "Txz"), (it + 1.0) * self.dt, True)

    def post_velocity(self, it):
        self._ring(("Txx", "Tzz", "Txz"), (it + 1.0) * self.dt, False)
        r = self.r
        full = np.full((self.bz, self.bx), self.b)
        self.rvel(ctypes.c_int(self.bx), ctypes.c_int(self.bz), _p(r["Txx"]), _p(r["Tzz"]), _p(r["Txz"]),
                  _p(r["U"]), _p(r["W"]), _p(full), _p(full))


# the FP64 tree's coefficient rows by order (PSV/psv_clean_double/param/const.h:13-18)
G_COEF = np.array([[1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], [9. / 8., 1. / 24., 0.0, 0.0],
                   [75. / 64., 25. / 384., 3. / 640., 0.0], [1225. / 1024., 245. / 3072., 49. / 5120., 5. / 7168.]])


class FileBox:
    """The reference's own box: a bw x bw square around (xs, zs) driven by
    the displacement frames of its sourcefile.

    Geometry and rings: BOX::init_box (PSV/psv_clean_double/box/box.cpp:
    10-95), box cell (i, j) at global (xs + s0 + i, zs + s0 + j) with
    s0 = (bw+1)/2 + 1 - bw.  Drive: BOX::setboxsource (:120-197), step it
    from frames it (old) and it+1 (new), each (U, W) on (bw+8)^2 cells.
    Box medium: MATERIAL::init_for_box_from_full (material/material.cpp:
    39-99).  frames: (nframes, 2, bw+8, bw+8) float64."""

    NAMES = ("U", "W", "Txx", "Tzz", "Txz")

    def __init__(self, sim, xs, zs, bw, frames, srcpvel, srcsvel, srcden, h, dt):
        s0 = (bw + 1) // 2 + 1 - bw
        self.sim, self.bw, self.dt = sim, bw, dt
        self.x0, self.z0, self.bx, self.bz = xs + s0, zs + s0, bw, bw
        self.frames = np.ascontiguousarray(frames, np.float64)
        assert self.frames.shape[1:] == (2, bw + 8, bw + 8)
        mu = srcsvel * srcsvel * srcden
        lam = (srcpvel * srcpvel * srcden - 2.0 * mu)
        b = 1.0 / srcden
        mu = mu * dt / h
        lam = lam * dt / h
        b = b * dt / h
        sl = (slice(self.z0, self.z0 + bw), slice(self.x0, self.x0 + bw))
        m = {k: sim.mat[k][sl] for k in ("BU", "BW", "MU", "MUA", "LAM")}
        g = {k: v.astype(np.float64) for k, v in m.items()}
        eps = 1e-4
        bad = ((np.abs(b - g["BU"]) > b * eps) | (np.abs(b - g["BW"]) > b * eps) |
               (np.abs(mu - g["MU"]) > mu * eps) | (np.abs(mu - g["MUA"]) > mu * eps) |
               (np.abs(lam - g["LAM"]) > lam * eps))
        for k, v in (("BU", b), ("BW", b), ("MU", mu), ("MUA", mu), ("LAM", lam)):
            m[k][bad] = v
        self.nforced = int(bad.sum())
        self.bmat = {k: np.ascontiguousarray(sim.mat[k][sl]) for k in m}
        dt_ = sim.dtype
        self.r = {k: np.zeros((bw, bw), dt_) for k in self.NAMES}
        j, i = np.mgrid[0:bw, 0:bw]
        d = np.minimum(np.minimum(i, j), np.minimum(bw - 1 - i, bw - 1 - j))
        self.sub, self.add = (d <= 3), (d >= 4) & (d <= 7)
        # derivative order of each cell (box.cpp:151-152), tx = i + 4, ty = j + 4
        mid = (bw + 1) // 2
        ordr = np.minimum(4, np.maximum(np.abs(j + 4 - 3 - mid) - 1, np.abs(i + 4 - 3 - mid) - 1))
        self.coef = G_COEF[np.maximum(ordr, 0)]
        self.rstress = getattr(lib(), "or_box_stress" + _sfx(dt_))
        self.rvel = getattr(lib(), "or_box_velocity" + _sfx(dt_))
        self.rstress.restype = self.rvel.restype = None
        self.s = None

    def source(self, it):
        """The five drive arrays of step it (bw x bw, float64), box.cpp:142-196."""
        bw, dt = self.bw, self.dt
        un, wn = self.frames[it + 1]
        uo, wo = self.frames[it]
        c0, c1, c2, c3 = (self.coef[..., k] for k in range(4))

        def sh(f, dx, dz):  # f(tx + dx, ty + dz) over the box cells
            return f[4 + dz:4 + dz + bw, 4 + dx:4 + dx + bw]

        ux = c0 * (sh(un, 0, 0) - sh(un, -1, 0))
        ux = ux - c1 * (sh(un, 1, 0) - sh(un, -2, 0))
        ux = ux + c2 * (sh(un, 2, 0) - sh(un, -3, 0))
        ux = ux - c3 * (sh(un, 3, 0) - sh(un, -4, 0))
        uz = c0 * (sh(un, 0, 0) - sh(un, 0, -1))
        uz = uz - c1 * (sh(un, 0, 1) - sh(un, 0, -2))
        uz = uz + c2 * (sh(un, 0, 2) - sh(un, 0, -3))
        uz = uz - c3 * (sh(un, 0, 3) - sh(un, 0, -4))
        wx = c0 * (sh(wn, 1, 0) - sh(wn, 0, 0))
        wx = wx - c1 * (sh(wn, 2, 0) - sh(wn, -1, 0))
        wx = wx + c2 * (sh(wn, 3, 0) - sh(wn, -2, 0))
        wx = wx - c3 * (sh(wn, 4, 0) - sh(wn, -3, 0))
        wz = c0 * (sh(wn, 0, 1) - sh(wn, 0, 0))
        wz = wz - c1 * (sh(wn, 0, 2) - sh(wn, 0, -1))
        wz = wz + c2 * (sh(wn, 0, 3) - sh(wn, 0, -2))
        wz = wz - c3 * (sh(wn, 0, 4) - sh(wn, 0, -3))
        lam = self.bmat["LAM"].astype(np.float64)
        mu = self.bmat["MU"].astype(np.float64)
        mua = self.bmat["MUA"].astype(np.float64)
        gam = lam + 2.0 * mu
        return {"U": (sh(un, 0, 0) - sh(uo, 0, 0)) / dt, "W": (sh(wn, 0, 0) - sh(wo, 0, 0)) / dt,
                "Txx": (gam * ux + lam * wz) / dt, "Tzz": (lam * ux + gam * wz) / dt,
                "Txz": (mua * (wx + uz)) / dt}

    def _ring(self, names, add):
        sl = (slice(self.z0, self.z0 + self.bz), slice(self.x0, self.x0 + self.bx))
        m = self.add if add else self.sub
        for nm in names:
            s = self.s[nm].astype(self.sim.dtype)
            a = getattr(self.sim, nm)[sl]
            if add:
                a[m] = self.r[nm][m] + s[m]
            else:
                self.r[nm][m] = a[m] - s[m]

    def pre_stress(self, it):
        self.s = self.source(it)
        self._ring(("U", "W"), True)

    def post_stress(self, it):
        self._ring(("U", "W"), False)
        r, bm = self.r, self.bmat
        self.rstress(ctypes.c_int(self.bx), ctypes.c_int(self.bz), _p(r["U"]), _p(r["W"]), _p(r["Txx"]),
                     _p(r["Tzz"]), _p(r["Txz"]), _p(bm["LAM"]), _p(bm["MU"]), _p(bm["MUA"]))
        self._ring(("Txx", "Tzz", "Txz"), True)

    def post_velocity(self, it):
        self._ring(("Txx", "Tzz", "Txz"), False)
        r, bm = self.r, self.bmat
        self.rvel(ctypes.c_int(self.bx), ctypes.c_int(self.bz), _p(r["Txx"]), _p(r["Tzz"]), _p(r["Txz"]),
                  _p(r["U"]), _p(r["W"]), _p(bm["BU"]), _p(bm["BW"]))


def run_psv_box(sim, box, nt, rx=None, rz=None, it0=0):
    """Box loop of PSV/psv_clean_double/main.cpp:153-200 (no point source)."""
    su, sw = [], []
    for it in range(it0, it0 + nt):
        box.pre_stress(it)
        sim.stress(it)
        box.post_stress(it)
        sim.velocity(it)
        if rx is not None:
            su.append(sim.record("U", rx, rz))
            sw.append(sim.record("W", rx, rz))
        box.post_velocity(it)
    return (np.array(su), np.array(sw)) if rx is not None else (None, None)
