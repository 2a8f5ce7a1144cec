"""Python side of the serving engine (include/fd23_serve.h).

``Server`` wraps the native engine: ``submit`` hands a ``Solve`` to it
and returns at once, the engine's own thread admits, batches, steps and
retires the work, and ``result`` returns the seismograms with the
request's latency.  ``serving_stream`` draws the mixed SH/P-SV request
stream (Poisson arrivals, per-request layered+GRF models, a quarter of the
requests with a plane-wave box); ``replay`` submits a stream on its
arrival schedule and collects latencies and throughput.
"""
from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, host
from ._lib import check

SIGMA_C = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168  # sum of |8th-order coefficients|


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


@dataclass
class Solve:
    """One request: a model, its shots, receivers and accuracy target."""
    wave: int
    storage: int
    nx: int
    nz: int
    npml: int
    nt: int
    dx: float
    dt: float
    material: list                 # scaled arrays (fd23_set_material order), or None with model
    cpml: np.ndarray               # fd23_cpml_pack layout
    xs: np.ndarray
    zs: np.ndarray
    stf: np.ndarray
    M: np.ndarray
    rx: np.ndarray
    rz: np.ndarray
    src_kind: int = _lib.SRC_LINE
    topabsorb: bool = False
    itrecord: int = 1
    box: dict | None = None        # x0 z0 bx bz incident angle f0 t0 amp vp vs rho
    model_key: int = 0
    tag: str = ""
    model: list | None = None      # vp, vs, den: the engine computes the material on the device

    @property
    def nshots(self):
        return len(self.xs)

    @property
    def ncomp(self):
        return 2 if self.wave == _lib.WAVE_PSV else 1

    @property
    def cells(self):
        return float(self.nshots) * self.nx * self.nz * self.nt


class Server:
    def __init__(self, hbm_budget=0, max_active=8, chunk_steps=256, max_batch_shots=64, autotune=False,
                 spare_shots=0):
        self.lib = _lib.load()
        d = _lib.ServerDesc(int(hbm_budget), max_active, chunk_steps, max_batch_shots, int(bool(autotune)),
                            int(spare_shots))
        h = ctypes.c_void_p()
        check(self.lib.fd23_server_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self._meta = {}
        self._lock = threading.Lock()

    def close(self):
        if getattr(self, "h", None):
            self.lib.fd23_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, s: Solve) -> int:
        f64 = s.storage == _lib.STORE_F64
        mdt = np.float64 if f64 else np.float32
        if s.material is not None:
            mats = [np.ascontiguousarray(m, mdt) for m in s.material]
            ptrs = ctypes.cast((ctypes.c_void_p * len(mats))(*[m.ctypes.data for m in mats]),
                               ctypes.POINTER(ctypes.c_void_p))
            mptrs = None
        else:
            mats = [None if m is None else np.ascontiguousarray(m, mdt) for m in s.model]
            mptrs = ctypes.cast((ctypes.c_void_p * 3)(*[None if m is None else m.ctypes.data for m in mats]),
                                ctypes.POINTER(ctypes.c_void_p))
            ptrs = None
        cp = np.ascontiguousarray(s.cpml, np.float64)
        xs = np.ascontiguousarray(s.xs, np.int32)
        zs = np.ascontiguousarray(s.zs, np.int32)
        stf = np.ascontiguousarray(s.stf, np.float64)
        M = np.ascontiguousarray(s.M, np.float64)
        rx = np.ascontiguousarray(s.rx, np.int32)
        rz = np.ascontiguousarray(s.rz, np.int32)
        b = s.box or {}
        r = _lib.Request(
            s.wave, s.storage, s.nx, s.nz, s.npml, int(bool(s.topabsorb)), s.nshots, s.nt, s.itrecord,
            s.dx, s.dt, ptrs, _vp(cp), s.src_kind, _vp(xs), _vp(zs),
            _vp(stf), stf.size, _vp(M), rx.size, _vp(rx), _vp(rz), int(bool(s.box)), b.get("x0", 0),
            b.get("z0", 0), b.get("bx", 0), b.get("bz", 0), b.get("incident", 0), b.get("angle", 0.0),
            b.get("f0", 0.0), b.get("t0", 0.0), b.get("amp", 0.0), b.get("vp", 0.0), b.get("vs", 0.0),
            b.get("rho", 0.0), int(s.model_key), mptrs)
        rid = ctypes.c_longlong()
        check(self.lib.fd23_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        # FP16/BF16 requests with a box run at FP32 storage
        eff = _lib.STORE_F32 if (s.box and s.storage in (_lib.STORE_F16, _lib.STORE_BF16)) else s.storage
        with self._lock:
            self._meta[rid.value] = (s.nshots, s.ncomp, -(-s.nt // s.itrecord), rx.size,
                                     np.float64 if eff == _lib.STORE_F64 else np.float32)
        return rid.value

    def wait(self, rid, timeout=-1.0):
        st = ctypes.c_int()
        check(self.lib.fd23_server_wait(self.h, rid, float(timeout), ctypes.byref(st)))
        return st.value

    def result(self, rid):
        """(seismograms [shot][comp][frame][rec], latency_s, queue_s)."""
        with self._lock:
            ns, nc, nf, nr, dt = self._meta.pop(rid)
        out = np.zeros((ns, nc, nf, nr), dt)
        lat, que = ctypes.c_double(), ctypes.c_double()
        code = self.lib.fd23_server_result(self.h, rid, ctypes.c_void_p(out.ctypes.data), out.nbytes,
                                           ctypes.byref(lat), ctypes.byref(que))
        check(code)
        return out, lat.value, que.value

    def stats(self):
        s = _lib.ServerStats()
        check(self.lib.fd23_server_stats_get(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


# ---------------------------------------------------------------------------
# request stream
# ---------------------------------------------------------------------------
GRIDS = ((256, 128), (512, 256), (1024, 512), (2048, 1024))


def make_solve(wave, storage, nx, nz, dx, nt, nshots, seed, npml=20, f0=10.0, box=None, model_key=None,
               cfl=0.6, tag="", device_material=False):
    """A request on a layered+GRF model (the batched-config recipe) with a
    CFL-safe time step dt = cfl*dx/(vpmax*sum|c|*sqrt 2), Ricker f0 sources
    at z=8 spread between the C-PML strips, receiver pairs at z=4.  With
    device_material the request carries vp, vs, den and the engine computes
    the material on the device."""
    from .workloads import layered_grf_model

    vp, vs, den = layered_grf_model(nx, nz, dx, seed)
    vpmax = float(vp.max())
    dt = cfl * dx / (vpmax * SIGMA_C * np.sqrt(2.0))
    mat = None
    if wave == _lib.WAVE_PSV:
        if not device_material:
            mat, _ = host.psv_material(vp, vs, den, dx, dt)
        M = host.psv_moment(_lib.SRC_LINE, 0.0, 0.0, 0.0, 0.0, dx, dt, explosion=True)
        kind = _lib.SRC_LINE
    else:
        if not device_material:
            mat, _ = host.sh_material(vs, den, dx, dt)
        M = host.sh_moment(_lib.SRC_LINE, 0.0, 90.0, 0.0, 0.0, dx, dt)
        kind = _lib.SRC_FORCE
    packed = host.cpml_pack(wave, nx, nz, npml, dt, 1e-6, vpmax / dx, f0)
    stf = host.ricker(nt, dt, f0)
    lo, hi = npml + 1, nx - npml - 2
    xs = np.array([lo + (k * (hi - lo)) // max(nshots - 1, 1) for k in range(nshots)], np.int32)
    zs = np.full(nshots, 8, np.int32)
    cols = np.arange(1, nx, 2, dtype=np.int32)
    rx = np.empty(2 * cols.size, np.int32)
    rx[0::2], rx[1::2] = cols, cols - 1
    rz = np.full(rx.size, 4, np.int32)
    if box is not None:
        x0, z0, bx, bz = box["x0"], box["z0"], box["bx"], box["bz"]
        sl = (slice(z0, z0 + bz), slice(x0, x0 + bx))
        box = dict(box, vp=float(vp[sl].mean()), vs=float(vs[sl].mean()), rho=float(den[sl].mean()))
        box.setdefault("t0", 1.2 / box["f0"])
        box.setdefault("amp", 1.0)
    model = [vp if wave == _lib.WAVE_PSV else None, vs, den] if device_material else None
    return Solve(wave, storage, nx, nz, npml, nt, dx, dt, None if mat is None else list(mat), packed, xs, zs, stf,
                 M, rx, rz, kind, box=box, model_key=seed if model_key is None else model_key, tag=tag, model=model)


def serving_stream(n, seed=5005, rate=40.0, storages=(_lib.STORE_F32, _lib.STORE_F16), models=0):
    """n requests: Poisson arrivals (rate per second), grids uniform over
    GRIDS, dx U(5,10) m, nt U{1000..8000}, 1-16 shots, SH or P-SV with
    p=0.5, FP32 or FP16 target with p=0.5, and a 200x100-cell plane-wave
    box (P or SV, Ricker 10 Hz, incidence U(0,30) degrees) on half of the
    P-SV requests (a quarter of all).  Model seed 5005 + request id.

    models > 0: a survey-like stream instead, where every request is a shot
    gather on one of `models` fixed problems (model, grid, physics, step
    count and target drawn as above, seed 5005 + problem id) with its own
    number of shots and shot positions, so requests on the same problem
    may share a launch."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        if models > 0:
            k = int(rng.integers(models))
            r = np.random.default_rng(seed + 1 + k)
        else:
            r = np.random.default_rng(seed + 1 + i)
        nx, nz = GRIDS[r.integers(len(GRIDS))]
        wave = _lib.WAVE_PSV if r.random() < 0.5 else _lib.WAVE_SH
        storage = storages[r.integers(len(storages))]
        dx = float(r.uniform(5.0, 10.0))
        nt = int(r.integers(1000, 8001))
        nshots = int(r.integers(1, 17))
        box = None
        if wave == _lib.WAVE_PSV and r.random() < 0.5:
            bx, bz = 200, 100
            box = dict(x0=(nx - bx) // 2, z0=max(4, (nz - 20 - bz) // 2), bx=bx, bz=bz,
                       incident=int(r.integers(2)), angle=float(r.uniform(0.0, 30.0)), f0=10.0)
        if models > 0:
            nshots = int(rng.integers(1, 17))
            s = make_solve(wave, storage, nx, nz, dx, nt, nshots, seed + 1 + k, box=box, tag="req%d" % i)
            # this gather's shot positions: shifted by up to +-8 cells, kept
            # between the C-PML strips
            shift = int(rng.integers(-8, 9))
            s.xs = np.clip(s.xs + shift, s.npml + 1, s.nx - s.npml - 2).astype(np.int32)
        else:
            s = make_solve(wave, storage, nx, nz, dx, nt, nshots, seed + 1 + i, box=box, tag="req%d" % i)
        out.append((float(t[i]), s))
    return out


def replay(server, stream, keep=False, on_result=None):
    """Submit (arrival_s, Solve) pairs on their schedule; a collector thread
    fetches results as they complete.  Returns per-request records and the
    wall time from the first arrival to the last completion."""
    recs = {}
    ids = []
    done = threading.Event()
    lock = threading.Lock()

    def collect():
        pending = []
        while True:
            with lock:
                pending.extend(ids)
                ids.clear()
            if not pending and done.is_set():
                return
            still = []
            for rid in pending:
                st = server.wait(rid, 0.0)
                if st in (_lib.REQ_DONE, _lib.REQ_FAILED):
                    try:
                        seis, lat, que = server.result(rid)
                        rec = recs[rid]
                        rec.update(latency=lat, queue=que, ok=True,
                                   norm=float(np.linalg.norm(seis.astype(np.float64))),
                                   finite=bool(np.isfinite(seis).all()))
                        if keep:
                            rec["seis"] = seis
                        if on_result:
                            on_result(rec, seis)
                    except _lib.Fd23Error as e:
                        recs[rid].update(ok=False, error=str(e))
                else:
                    still.append(rid)
            pending = still
            if pending:
                time.sleep(0.0005)
            else:
                time.sleep(0.001)

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for at, s in stream:
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        rid = server.submit(s)
        recs[rid] = dict(tag=s.tag, wave=s.wave, storage=s.storage, nx=s.nx, nz=s.nz, nt=s.nt,
                         nshots=s.nshots, box=bool(s.box), cells=s.cells, arrival=at)
        with lock:
            ids.append(rid)
    done.set()
    th.join()
    wall = time.perf_counter() - t0
    return list(recs.values()), wall
