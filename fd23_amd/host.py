"""Python access to the library's host precompute (C-PML, material, sources,
models, ISIS, par files).  All arithmetic happens in the C++ library."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def cpml_profile(n, npml, startpos, pml_dt, pml_r, pml_v, pml_fc):
    b, c, k = (np.zeros(n, np.float64) for _ in range(3))
    check(_lib.load().fd23_cpml_profile(n, npml, startpos, pml_dt, pml_r, pml_v, pml_fc, _ptr(b), _ptr(c),
                                        _ptr(k)))
    return b, c, k


def cpml_pack(wave, nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc):
    L = _lib.load()
    out = np.zeros(L.fd23_cpml_pack_len(wave, nx, nz), np.float64)
    check(L.fd23_cpml_pack(wave, nx, nz, npml, pml_dt, pml_r, pml_v, pml_fc, _ptr(out)))
    return out


def psv_material(vp, vs, den, h, dt, f64=False):
    """Returns ([BU, BW, MU, MUA, LAM], stab); raises on stab > 1."""
    dt_ = np.float64 if f64 else np.float32
    vp, vs, den = (np.ascontiguousarray(a, dt_) for a in (vp, vs, den))
    nz, nx = vp.shape
    out = [np.zeros((nz, nx), dt_) for _ in range(5)]
    stab = ctypes.c_double()
    f = _lib.load().fd23_psv_material_f64 if f64 else _lib.load().fd23_psv_material
    check(f(nx, nz, h, dt, _ptr(vp), _ptr(vs), _ptr(den), *[_ptr(a) for a in out], ctypes.byref(stab)))
    return out, stab.value


def sh_material(vs, den, h, dt):
    vs, den = (np.ascontiguousarray(a, np.float32) for a in (vs, den))
    nz, nx = vs.shape
    out = [np.zeros((nz, nx), np.float32) for _ in range(3)]
    stab = ctypes.c_double()
    check(_lib.load().fd23_sh_material(nx, nz, h, dt, _ptr(vs), _ptr(den), *[_ptr(a) for a in out],
                                       ctypes.byref(stab)))
    return out, stab.value


def psv_moment(kind, strike, dip, rake, azimuth, h, dt, explosion=False):
    M = np.zeros(3, np.float64)
    check(_lib.load().fd23_psv_moment(kind, int(bool(explosion)), strike, dip, rake, azimuth, h, dt, _ptr(M)))
    return M


def sh_moment(kind, strike, dip, rake, azimuth, h, dt):
    M = np.zeros(2, np.float64)
    check(_lib.load().fd23_sh_moment(kind, strike, dip, rake, azimuth, h, dt, _ptr(M)))
    return M


def gaussian(alpha):
    a = int(alpha)
    cap = 2 * abs(int(a * 3.0)) + 2
    out = np.zeros(cap, np.float64)
    n = ctypes.c_int()
    check(_lib.load().fd23_stf_gaussian(a, _ptr(out), cap, ctypes.byref(n)))
    return out[:n.value].copy()


def trapezoid(dt, t1, t2, t3):
    cap = int(t1 / dt) + int(t2 / dt) + int(t3 / dt) + 8
    out = np.zeros(cap, np.float64)
    n = ctypes.c_int()
    check(_lib.load().fd23_stf_trapezoid(dt, t1, t2, t3, _ptr(out), cap, ctypes.byref(n)))
    return out[:n.value].copy()


def ricker(n, dt, f0, t0=None):
    if t0 is None:
        t0 = 1.2 / f0
    out = np.zeros(n, np.float64)
    check(_lib.load().fd23_stf_ricker(n, dt, f0, t0, _ptr(out)))
    return out


def model_read(name, nx, nz, h):
    vp, vs, den = (np.zeros((nz, nx), np.float32) for _ in range(3))
    check(_lib.load().fd23_model_read(name.encode(), nx, nz, h, _ptr(vp), _ptr(vs), _ptr(den)))
    return vp, vs, den


def isis_write(path, raw, rx, rz, shiftx, shiftz, h, dt, itrecord, xs, zs):
    raw = np.ascontiguousarray(raw, np.float32)
    rx = np.ascontiguousarray(rx, np.int32)
    rz = np.ascontiguousarray(rz, np.int32)
    ntrec, nrec = raw.shape
    check(_lib.load().fd23_isis_write(path.encode(), _ptr(raw), nrec, ntrec, _ptr(rx), _ptr(rz), shiftx, shiftz,
                                      h, dt, itrecord, xs, zs))


def psv_moment_f64(kind, strike, dip, rake, azimuth, h, dt):
    """The FP64 tree's moment tensor (double throughout)."""
    M = np.zeros(3, np.float64)
    check(_lib.load().fd23_psv_moment_f64(kind, strike, dip, rake, azimuth, h, dt, _ptr(M)))
    return M


def gaussian_f64(alpha):
    a = int(alpha)
    cap = 2 * abs(int(a * 3.0)) + 2
    out = np.zeros(cap, np.float64)
    n = ctypes.c_int()
    check(_lib.load().fd23_stf_gaussian_f64(a, _ptr(out), cap, ctypes.byref(n)))
    return out[:n.value].copy()


def trapezoid_f64(dt, t1, t2, t3):
    cap = int(t1 / dt) + int(t2 / dt) + int(t3 / dt) + 8
    out = np.zeros(cap, np.float64)
    n = ctypes.c_int()
    check(_lib.load().fd23_stf_trapezoid_f64(dt, t1, t2, t3, _ptr(out), cap, ctypes.byref(n)))
    return out[:n.value].copy()


def model_read_f64(name, nx, nz, h):
    vp, vs, den = (np.zeros((nz, nx), np.float64) for _ in range(3))
    check(_lib.load().fd23_model_read_f64(name.encode(), nx, nz, h, _ptr(vp), _ptr(vs), _ptr(den)))
    return vp, vs, den


def isis_write_f64(path, raw, rx, rz, shiftx, shiftz, h, dt, itrecord, xs, zs, src_shift_x=0.25,
                   src_shift_z=-0.25, bw=0):
    raw = np.ascontiguousarray(raw, np.float64)
    rx = np.ascontiguousarray(rx, np.int32)
    rz = np.ascontiguousarray(rz, np.int32)
    ntrec, nrec = raw.shape
    check(_lib.load().fd23_isis_write_f64(path.encode(), _ptr(raw), nrec, ntrec, _ptr(rx), _ptr(rz), shiftx,
                                          shiftz, h, dt, itrecord, xs, zs, src_shift_x, src_shift_z, bw))


def line_receivers(nsta, ix0, iz0, idx, idz):
    rx = np.zeros(2 * nsta, np.int32)
    rz = np.zeros(2 * nsta, np.int32)
    check(_lib.load().fd23_line_receivers(nsta, ix0, iz0, idx, idz, _ptr(rx), _ptr(rz)))
    return rx, rz


def station_read(path):
    lib = _lib.load()
    n = ctypes.c_int()
    check(lib.fd23_station_read(path.encode(), None, None, 0, ctypes.byref(n)))
    rx = np.zeros(n.value, np.int32)
    rz = np.zeros(n.value, np.int32)
    check(lib.fd23_station_read(path.encode(), _ptr(rx), _ptr(rz), n.value, ctypes.byref(n)))
    return rx, rz


class Par:
    """getpar-compatible parameters: ``Par(["prog", "par=file", "nx=100"])``."""

    def __init__(self, argv):
        self.lib = _lib.load()
        arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
        h = ctypes.c_void_p()
        check(self.lib.fd23_par_open(len(argv), arr, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.fd23_par_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, fn, name, ctype):
        v = ctype()
        r = fn(self.h, name.encode(), ctypes.byref(v))
        if r < 0:
            check(r)
        return v.value if r == 1 else None

    def int(self, name, default=None):
        v = self._get(self.lib.fd23_par_int, name, ctypes.c_int)
        return default if v is None else v

    def float(self, name, default=None):
        v = self._get(self.lib.fd23_par_float, name, ctypes.c_float)
        return default if v is None else v

    def double(self, name, default=None):
        v = self._get(self.lib.fd23_par_double, name, ctypes.c_double)
        return default if v is None else v

    def str(self, name, default=None):
        buf = ctypes.create_string_buffer(4096)
        r = self.lib.fd23_par_string(self.h, name.encode(), buf, 4096)
        if r < 0:
            check(r)
        return buf.value.decode() if r == 1 else default

    def must(self, kind, name):
        v = getattr(self, kind)(name)
        if v is None:
            raise KeyError("must specify value for '%s'" % name)
        return v


def punch_blocks(times, before, after, dt, f64=False):
    """Per-block first/last step tables of the first-arrival window
    (fd23_punch_blocks, or fd23_punch_blocks_f64 for the FP64 tree)."""
    t = np.ascontiguousarray(times, np.float64 if f64 else np.float32)
    nz, nx = t.shape
    first = np.zeros((nz // 16, nx // 32), np.int32)
    last = np.zeros_like(first)
    fn = _lib.load().fd23_punch_blocks_f64 if f64 else _lib.load().fd23_punch_blocks
    check(fn(_ptr(t), nx, nz, before, after, dt, _ptr(first), _ptr(last)))
    return first, last
