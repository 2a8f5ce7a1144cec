"""Host-side mirror of the reference's per-GPU objects on top of the C ABI.

``Context`` owns one device context: the FIELD, MATERIAL, CPML and RECORD
state that the reference's ``singlegpu()`` builds
(PSV/clean_nobox/main.cpp:41-81), for ``nshots`` shots sharing one model.
Its stepping methods are the reference's per-step calls:

  reference (PSV/clean_nobox/main.cpp:115-150)   Context
  cu_step_stress(...)                             step_stress()
  cu_point_source / cu_point_source_p(...)        inject(it)
  cu_step_velocity(...)                           step_velocity()
  RECORD::cu_record(...)                          record(it, itrecord)
  the whole loop                                  run(it0, nt, itrecord)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check

PSV_FIELDS = ("U", "W", "Txx", "Txz", "Tzz")  # state%06d order
SH_FIELDS = ("V", "Tx", "Tz")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    def __init__(self, wave, nx, nz, npml, nshots=1, storage=_lib.STORE_F32, topabsorb=False,
                 max_frames=0, max_rec=0, stream=None):
        self.lib = _lib.load()
        self.wave, self.nx, self.nz, self.npml = wave, nx, nz, npml
        self.nshots, self.storage = nshots, storage
        self.topabsorb = bool(topabsorb)
        self.max_frames, self.max_rec = max_frames, max_rec
        self.cdtype = np.float64 if storage == _lib.STORE_F64 else np.float32
        d = _lib.Desc(wave, storage, nx, nz, npml, int(self.topabsorb), nshots, max_frames, max_rec,
                      stream)
        h = ctypes.c_void_p()
        check(self.lib.fd23_ctx_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.ncomp = 2 if wave == _lib.WAVE_PSV else 1
        self.nrec = np.zeros((nshots, self.ncomp), np.int64)

    # -- lifetime ------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.fd23_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- set-up ----------------------------------------------------------
    def set_material(self, arrays):
        """P-SV: BU, BW, MU, MUA, LAM; SH: BV, MUX, MUZ (nz x nx, scaled by dt/h)."""
        keep = [np.ascontiguousarray(a, self.cdtype) for a in arrays]
        for a in keep:
            assert a.shape == (self.nz, self.nx), a.shape
        ptrs = (ctypes.c_void_p * len(keep))(*[a.ctypes.data for a in keep])
        check(self.lib.fd23_set_material(self.h, ptrs))

    def set_model(self, vp, vs, den, h, dt, noalloc=False):
        """Material computed on the device from vp, vs, den (nz x nx); returns
        the stability number (raises Fd23Error above 1).  SH ignores vp.
        noalloc=True goes through fd23_set_model_noalloc (no device
        allocation; the library's fixed scratch)."""
        arrs = [None if a is None else np.ascontiguousarray(a, self.cdtype) for a in (vp, vs, den)]
        for a in arrs:
            assert a is None or a.shape == (self.nz, self.nx), a.shape
        stab = ctypes.c_double(0.0)
        p = [None if a is None else a.ctypes.data for a in arrs]
        fn = self.lib.fd23_set_model_noalloc if noalloc else self.lib.fd23_set_model
        check(fn(self.h, p[0], p[1], p[2], h, dt, ctypes.byref(stab)))
        return stab.value

    def set_model_file(self, path, h, dt):
        """A .mdl (rasterised on the device) or raster model file and its
        material computed on the device; returns the stability number."""
        stab = ctypes.c_double(0.0)
        check(self.lib.fd23_set_model_file(self.h, str(path).encode(), h, dt, ctypes.byref(stab)))
        return stab.value

    def material(self, index):
        """One material array (fd23_set_material order) as nz x nx."""
        out = np.zeros((self.nz, self.nx), self.cdtype)
        check(self.lib.fd23_get_material(self.h, index, _ptr(out)))
        return out

    def set_cpml(self, packed):
        packed = np.ascontiguousarray(packed, np.float64)
        n = self.lib.fd23_cpml_pack_len(self.wave, self.nx, self.nz)
        assert packed.size == n, (packed.size, n)
        check(self.lib.fd23_set_cpml(self.h, _ptr(packed)))

    def set_source(self, shot, kind, xs, zs, stf, M):
        stf = np.ascontiguousarray(stf, np.float64)
        M = np.ascontiguousarray(M, np.float64)
        check(self.lib.fd23_set_source(self.h, shot, kind, xs, zs, _ptr(stf), stf.size, _ptr(M)))

    def set_receivers(self, shot, comp, rx, rz):
        rx = np.ascontiguousarray(rx, np.int32)
        rz = np.ascontiguousarray(rz, np.int32)
        check(self.lib.fd23_set_receivers(self.h, shot, comp, _ptr(rx), _ptr(rz), rx.size))
        self.nrec[shot, comp] = rx.size

    def set_shot_slot(self, shot, active=True, t0=0):
        """Start shot `shot` at step t0 (its samples and frames count from
        there), or make its slot idle (not stepped by run)."""
        check(self.lib.fd23_set_shot_slot(self.h, shot, 1 if active else 0, t0))

    def reset_shot(self, shot):
        """Zero one shot's fields, C-PML memory and seismogram entries."""
        check(self.lib.fd23_reset_shot(self.h, shot))

    def set_scales(self, scales):
        s = np.ascontiguousarray(scales, np.float32)
        check(self.lib.fd23_set_scales(self.h, _ptr(s), s.size))

    def get_scales(self, n):
        s = np.zeros(n, np.float32)
        a = np.zeros(n, np.uint32)
        check(self.lib.fd23_get_scales(self.h, _ptr(s), _ptr(a), n))
        return s, a.view(np.float32)

    # -- stepping --------------------------------------------------------
    def step_stress(self):
        check(self.lib.fd23_step_stress(self.h))

    def inject(self, it):
        check(self.lib.fd23_inject(self.h, it))

    def step_velocity(self):
        check(self.lib.fd23_step_velocity(self.h))

    def record(self, it, itrecord=1):
        check(self.lib.fd23_record(self.h, it, itrecord))

    def set_punch(self, times, before=0.0, after=0.0, dt=1.0, f64=False):
        """First-arrival window from an nz x nx travel-time grid (None
        removes it); see fd23_set_punch.  f64=True forms the windows in
        double from a double grid, as the FP64 tree (fd23_set_punch_f64)."""
        if times is None:
            check(self.lib.fd23_set_punch(self.h, None, 0.0, 0.0, 1.0))
            return
        t = np.ascontiguousarray(times, np.float64 if f64 else np.float32)
        if t.shape != (self.nz, self.nx):
            raise ValueError(f"travel times {t.shape} != (nz, nx) = {(self.nz, self.nx)}")
        fn = self.lib.fd23_set_punch_f64 if f64 else self.lib.fd23_set_punch
        check(fn(self.h, _ptr(t), before, after, dt))

    def prepare(self, nt, itrecord=1):
        """Capture the graphs run(*, nt, itrecord) replays, without running."""
        check(self.lib.fd23_prepare(self.h, nt, itrecord))

    def run(self, it0, nt, itrecord=1):
        check(self.lib.fd23_run(self.h, it0, nt, itrecord))

    def sync(self):
        check(self.lib.fd23_sync(self.h))

    # -- outputs ---------------------------------------------------------
    def seismograms(self, shot, comp, nframes):
        n = int(self.nrec[shot, comp])
        out = np.zeros((nframes, n), self.cdtype)
        if n and nframes:
            check(self.lib.fd23_get_seismograms(self.h, shot, comp, _ptr(out), nframes))
        return out

    def field(self, name, shot=0):
        names = PSV_FIELDS if self.wave == _lib.WAVE_PSV else SH_FIELDS
        out = np.zeros((self.nz, self.nx), self.cdtype)
        check(self.lib.fd23_get_field(self.h, shot, names.index(name), _ptr(out)))
        return out

    def set_field(self, name, arr, shot=0):
        names = PSV_FIELDS if self.wave == _lib.WAVE_PSV else SH_FIELDS
        a = np.ascontiguousarray(arr, self.cdtype)
        check(self.lib.fd23_set_field(self.h, shot, names.index(name), _ptr(a)))

    def state_bytes(self):
        return int(self.lib.fd23_state_bytes(self.h))

    def get_state(self, shot=0):
        n = self.state_bytes()
        buf = np.zeros(n, np.uint8)
        check(self.lib.fd23_get_state(self.h, shot, _ptr(buf), n))
        return buf.tobytes()

    def set_state(self, data, shot=0):
        buf = np.frombuffer(data, np.uint8).copy()
        check(self.lib.fd23_set_state(self.h, shot, _ptr(buf), buf.size))

    # -- box ---------------------------------------------------------------
    def box_configure(self, x0, z0, bx, bz, incident, angle_deg, f0, t0, amp, vp, vs, rho, h, dt):
        """Plane-wave box on every shot (incident 0 = P, 1 = SV)."""
        check(self.lib.fd23_box_configure(self.h, -1, x0, z0, bx, bz, incident, angle_deg, f0, t0, amp,
                                          vp, vs, rho, h, dt))
        self._box_shape = (bz, bx)

    def box_configure_frames(self, xs, zs, bw, frames, srcpvel, srcsvel, srcden, h, dt):
        """The reference's file-driven box (fd23_box_configure_frames):
        frames is (nframes, 2, bw+8, bw+8) float64 displacement (U, W)."""
        f = np.ascontiguousarray(frames, np.float64)
        assert f.ndim == 4 and f.shape[1:] == (2, bw + 8, bw + 8), f.shape
        check(self.lib.fd23_box_configure_frames(self.h, xs, zs, bw, _ptr(f), f.shape[0], srcpvel, srcsvel,
                                                 srcden, h, dt))
        self._box_shape = (bw, bw)
        return self.lib.fd23_box_forced_cells(self.h)

    def box_residual(self, shot=0):
        """Residual grid of the box: dict of U W Txx Tzz Txz (bz x bx)."""
        bz, bx = self._box_shape
        out = np.zeros((5, bz, bx), self.cdtype)
        check(self.lib.fd23_box_get_residual(self.h, shot, _ptr(out)))
        return dict(zip(("U", "W", "Txx", "Tzz", "Txz"), out))

    # -- tuning / measurement ----------------------------------------------
    @property
    def tile(self):
        return self.lib.fd23_get_tile(self.h)

    @tile.setter
    def tile(self, t):
        check(self.lib.fd23_set_tile(self.h, t))

    def tile_shape(self, tile=None):
        """(rz, spb, seg_rows, block_rows) of a tile index (default: the
        current one): seg_rows and block_rows are 0 for tiled kinds; the
        fused P-SV step has segments of seg_rows rows in blocks of
        block_rows (0: the tile-parallel step); the SH block step runs
        seg_rows steps per launch over regions of block_rows rows, rz rows
        per thread."""
        t = self.tile if tile is None else tile
        v = [ctypes.c_int() for _ in range(4)]
        check(self.lib.fd23_tile_shape(t, *(ctypes.byref(x) for x in v)))
        return tuple(x.value for x in v)

    def tile_kind(self, tile=None):
        """TILE_TILED or TILE_FUSED for a tile index."""
        k = self.lib.fd23_tile_kind(self.tile if tile is None else tile)
        if k < 0:
            check(k)
        return k

    def tile_split(self, tile=None):
        """True when the tile shape launches interior and edge tiles apart."""
        k = self.lib.fd23_tile_split(self.tile if tile is None else tile)
        if k < 0:
            check(k)
        return bool(k)

    @property
    def chains(self):
        """Number of independent chains of passes run() captures, each over
        a part of the shot groups (1: one chain)."""
        return self.lib.fd23_get_chains(self.h)

    @chains.setter
    def chains(self, n):
        check(self.lib.fd23_set_chains(self.h, int(n)))

    def autotune(self, steps=4):
        best = ctypes.c_double()
        check(self.lib.fd23_autotune(self.h, steps, ctypes.byref(best)))
        return self.tile, best.value

    def bytes_per_step(self):
        return float(self.lib.fd23_bytes_per_step(self.h))

    def device_bytes(self):
        return int(self.lib.fd23_ctx_device_bytes(self.h))

    def time_passes(self, nsteps):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.fd23_time_passes(self.h, nsteps, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def desc_device_bytes(wave, nx, nz, npml, nshots=1, storage=_lib.STORE_F32, topabsorb=False,
                      max_frames=0, max_rec=0):
    lib = _lib.load()
    d = _lib.Desc(wave, storage, nx, nz, npml, int(bool(topabsorb)), nshots, max_frames, max_rec, None)
    return int(lib.fd23_desc_device_bytes(ctypes.byref(d)))
