#!/usr/bin/env python3
"""Throughput benchmark of the fd23 staggered-grid loop on one MI355X.

A step is one time step of the loop (stress pass with fused C-PML, source
injection, velocity pass with fused C-PML, receiver gather) over every shot
of the workload, replayed from HIP graphs (captured before the timed
region).  Prints ONE JSON line.

Default workload (the largest single-GPU configuration): P-SV 4096x2048,
dx=4 m, dt=0.25 ms, npml 32, seeded 12-layer + GRF model, 8 shots of an
explosive Ricker 20 Hz at z=8 spread in x, 2048 receiver pairs per shot at
z=4 (vx, vz), FP32 storage.  Inputs are resident in HBM before timing.  The
same line carries the FP16 and BF16 storage rates of that workload with
their seismogram error against FP64 over a stated number of steps, and the
64-shot 1024x512 batch's and one 1024x512 shot's shots/s.

roofline: algorithmic HBM bytes of the dominant pass (fields read once,
updated arrays read and written once, material once, psi read+written in
the C-PML strips) divided by that pass's average duration from HIP events
on the context stream.  cpu_baseline: the oracle's C restatement of the
same passes on this host's cores over a bounded number of steps of one
shot of the workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def fused_bytes(wl):
    """Algorithmic bytes of one fused P-SV step: the five fields read once
    and written once, material once per launch, psi read and written in the
    C-PML cells."""
    from fd23_amd import _lib

    S, N = wl.nshots, float(wl.nx) * wl.nz
    wf = {_lib.STORE_F32: 4, _lib.STORE_F16: 2, _lib.STORE_BF16: 2, _lib.STORE_F64: 8}[wl.storage]
    wm = 8 if wl.storage == _lib.STORE_F64 else 4
    nxp = 2.0 * wl.npml * wl.nz
    nzp = (2.0 if wl.topabsorb else 1.0) * wl.npml * wl.nx
    return S * (10 * wf * N + 8 * wf * (nxp + nzp)) + 5 * wm * N


def block_bytes(wl, k):
    """Algorithmic bytes of one SH block launch of k steps: V, Tx and Tz
    read once and written once, the four psi read and written in the C-PML
    cells, the three material arrays read once (the recomputed halos are
    re-reads, not algorithmic bytes)."""
    from fd23_amd import _lib

    S, N = wl.nshots, float(wl.nx) * wl.nz
    wf = {_lib.STORE_F32: 4, _lib.STORE_F64: 8}[wl.storage]
    nxp = 2.0 * wl.npml * wl.nz
    nzp = (2.0 if wl.topabsorb else 1.0) * wl.npml * wl.nx
    return S * (6 * wf * N + 4 * wf * (nxp + nzp)) + 3 * wf * N


def pass_bytes(wl, ctx_bytes_per_step=None):
    """Algorithmic bytes of the stress and velocity passes of one step."""
    from fd23_amd import _lib

    S, N = wl.nshots, float(wl.nx) * wl.nz
    wf = {_lib.STORE_F32: 4, _lib.STORE_F16: 2, _lib.STORE_BF16: 2, _lib.STORE_F64: 8}[wl.storage]
    wm = 8 if wl.storage == _lib.STORE_F64 else 4
    nxp = 2.0 * wl.npml * wl.nz
    nzp = (2.0 if wl.topabsorb else 1.0) * wl.npml * wl.nx
    if wl.wave == _lib.WAVE_PSV:
        stress = S * (8 * wf * N + 4 * wf * (nxp + nzp)) + 3 * wm * N
        vel = S * (7 * wf * N + 4 * wf * (nxp + nzp)) + 2 * wm * N
    else:
        stress = S * (5 * wf * N + 2 * wf * (nxp + nzp)) + 2 * wm * N
        vel = S * (4 * wf * N + 2 * wf * (nxp + nzp)) + 1 * wm * N
    return stress, vel


def sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
            return
    except Exception:
        pass


def run_gpu(wl, steps, warmup, autotune=True, tile=None, passes=True, chains=None):
    from fd23_amd import _lib

    ctx = wl.context(max_frames=warmup + steps)
    tune_ms = None
    if tile is not None:
        ctx.tile = tile
        if chains is not None:
            ctx.chains = chains
    elif autotune:
        _, tune_ms = ctx.autotune(16)
    tile = ctx.tile
    chains = ctx.chains  # independent chains of passes over parts of the shots
    # the graphs of both runs are captured before any step runs
    ctx.prepare(warmup)
    ctx.prepare(steps)
    ctx.run(0, warmup, 1)
    ctx.sync()
    sync()
    t0 = time.perf_counter()
    ctx.run(warmup, steps, 1)
    ctx.sync()
    sync()
    el = time.perf_counter() - t0
    seis = ctx.seismograms(0, 0, warmup + steps)
    finite = bool(np.isfinite(seis).all())
    kind = ctx.tile_kind()
    fused = kind == _lib.TILE_FUSED and wl.wave == _lib.WAVE_PSV and wl.punch_after is None
    # steps per launch of the SH block step, 0 when another path runs
    block = ctx.tile_shape()[2] if kind == _lib.TILE_BLOCK and wl.punch_after is None else 0
    names = kernel_names(ctx, wl)
    s_ms, v_ms = ctx.time_passes(20) if passes else (None, None)
    ctx.close()
    return dict(elapsed=el, stress_ms=s_ms, vel_ms=v_ms, tile=tile, tune_ms=tune_ms, finite=finite, fused=fused,
                block=block, kernels=names, chains=chains)


STORE_TYPE = {0: "float", 1: "__half", 2: "__hip_bfloat16", 3: "double"}


def kernel_names(ctx, wl):
    """Kernels of the (stress, velocity) passes as profilers print them, a
    tuple per pass: split tile shapes launch the interior tiles and the edge
    tiles of a P-SV pass as two instances of one kernel; the fused step and
    the SH block step are one kernel for both."""
    from fd23_amd import _lib

    rz, spb, seg, blk = ctx.tile_shape()
    kind = ctx.tile_kind()
    t = STORE_TYPE[wl.storage]
    if wl.wave == _lib.WAVE_SH:
        if kind == _lib.TILE_BLOCK and wl.punch_after is None:
            n = ("fd23::k_sh_block<%s, %d, %d, %d>" % (t, seg, rz, blk // rz),)
            return (n, n)
        if kind != _lib.TILE_TILED:  # a window runs tile 0's tiled kernels
            rz, spb = ctx.tile_shape(0)[:2]
        return (("fd23::k_sh_stress<%s, %d, %d>" % (t, rz, spb),), ("fd23::k_sh_velocity<%s, %d, %d>" % (t, rz, spb),))
    # a first-arrival window runs the whole-grid tiled kernels (tile 0's
    # shape in place of a fused one)
    window = wl.punch_after is not None
    if kind == _lib.TILE_FUSED and not window:
        # block rows 0: the tile-parallel fused step, segments of 4*CZ rows
        n = ("fd23::k_psv_fused_tile<%s, %d>" % (t, seg // 4),) if blk == 0 else \
            ("fd23::k_psv_fused<%s, %d, %d>" % (t, seg, blk),)
        return (n, n)
    if kind == _lib.TILE_FUSED:
        rz, spb = ctx.tile_shape(0)[:2]
    regions = (1, 2) if ctx.tile_split() and not window else (0,)
    return tuple(tuple("fd23::k_psv_%s<%s, %d, %d, %d>" % (p, t, rz, spb, region) for region in regions)
                 for p in ("stress", "velocity"))


def measured_traffic(workload, kernels):
    """HBM bytes per pass (summed over its kernels) on a workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE summaries kept beside this file
    (scaled by the copy calibration stated there); None when a kernel of
    the pass was not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    total = 0
    for k in kernels:
        e = table.get("%s|%s" % (workload, k))
        if e is None:
            return None
        total += int(e["hbm_bytes"])
    return total


def cpu_baseline(wl, budget_s=12.0, budget_1t=4.0):
    """Oracle passes (C, OpenMP) on this host over a bounded sample: all
    cores, then one thread."""
    import ctypes

    import oracle as O
    from fd23_amd import _lib

    lib = O.lib()
    lib.or_num_threads.restype = ctypes.c_int
    lib.or_set_threads.restype = None
    cores = int(lib.or_num_threads())
    names = ("BU", "BW", "MU", "MUA", "LAM")
    mat = dict(zip(names, wl.mat))
    cp, off = {}, 0
    for name, n in zip(O.PSV_CPML_X + O.PSV_CPML_Z, [wl.nx] * 4 + [wl.nz] * 4):
        cp[name] = tuple(wl.packed[off + j * n: off + (j + 1) * n].astype(np.float32) for j in range(3))
        off += 3 * n
    assert wl.wave == _lib.WAVE_PSV
    sim = O.PSV(wl.nx, wl.nz, wl.npml, mat, cp, wl.topabsorb)
    stf, M = wl.stf.astype(np.float32), wl.M.astype(np.float32)

    def timed(budget, cap):
        sim.stress()
        sim.velocity()
        t0 = time.perf_counter()
        k = 0
        while True:
            sim.stress()
            sim.inject(0, k, stf, M, *wl.sources[0])
            sim.velocity()
            k += 1
            el = time.perf_counter() - t0
            if el > budget or k >= cap:
                return k, el

    k, el = timed(budget_s, 2000)
    lib.or_set_threads(1)
    k1, el1 = timed(budget_1t, 200)
    lib.or_set_threads(cores)
    shot = "" if wl.nshots == 1 else "shot 0 of the %d-shot workload, " % wl.nshots
    return dict(value=wl.nx * wl.nz * k / el / 1e9, unit="Gcell/s", cores=cores, kind="port",
                sample="%d steps of %sthe %dx%d P-SV FP32 loop (oracle C restatement, OpenMP, %d threads)"
                       % (k, shot, wl.nx, wl.nz, cores),
                value_1thread=wl.nx * wl.nz * k1 / el1 / 1e9, sample_1thread="%d steps, 1 thread" % k1,
                cpu=_cpu_model())


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


STORAGE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 3}
STORAGE_TAG = {0: "fp32", 1: "fp16", 2: "bf16", 3: "fp64"}


def workload_id(wl):
    """Workload name with the storage type it runs at, e.g.
    psv_4096x2048_x8_fp16 (the key of the profiles' traffic table)."""
    base = wl.name[:-5] if wl.name.endswith("_fp32") else wl.name
    return base + "_" + STORAGE_TAG[wl.storage]


def make_workload(name, storage=None):
    from fd23_amd import _lib, workloads as W

    if name == "psv2000":
        wl = W.config_psv_2000x1000()
    elif name == "batched":
        wl = W.config_psv_batched(64)
    elif name == "single1024":
        wl = W.config_psv_batched(1)
    elif name == "c4":
        wl = W.config_psv_precision(_lib.STORE_F32)
    elif name in ("f16", "bf16"):
        wl = W.config_psv_precision(STORAGE[name])
    elif name == "punch":
        wl = W.config_psv_punch()
    else:
        wl = W.config_sh_400x200()
    if storage is not None:
        wl.storage = STORAGE[storage]
    return wl


def seismograms(wl, nt):
    """All shots' and components' seismograms of nt steps, as float64."""
    ctx = wl.context(max_frames=nt)
    ctx.run(0, nt, 1)
    ctx.sync()
    out = np.stack([ctx.seismograms(s, c, nt).astype(np.float64) for s in range(wl.nshots)
                    for c in range(ctx.ncomp)])
    ctx.close()
    return out


def accuracy_vs_f64(name, storages, nt, built=None):
    """Seismograms of the workload at each storage type against an FP64
    storage run of the same inputs (FP64 on the GPU, pinned to the FP64
    oracle by the parity tests): relative L2 and peak error over all shots
    and components.  built: an already built FP32-material workload whose
    storage is switched (FP32, FP16 and BF16 share the FP32 material)."""
    ref = seismograms(make_workload(name, "f64").build(), nt)
    out = {}
    for st in storages:
        wl = built if built is not None else make_workload(name, st).build()
        keep = wl.storage
        wl.storage = STORAGE[st]
        a = seismograms(wl, nt)
        wl.storage = keep
        out[st] = {"steps": nt, "rel_l2": float(np.linalg.norm(a - ref) / np.linalg.norm(ref)),
                   "peak": float(np.abs(a - ref).max() / np.abs(ref).max()), "finite": bool(np.isfinite(a).all())}
    return out


def calibrate_copy(nbytes=1 << 30, reps=10):
    """Device copies of known size: torch's copy and the library's
    16-byte-lane streaming copy (fd23_device_copy) at several grid sizes,
    for the practical read+write bandwidth and for scaling HBM counters."""
    import ctypes

    import torch
    from fd23_amd import _lib

    lib = _lib.load()
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    out = {"copy_bytes_each_way": nbytes, "reps": reps, "torch_copy_GBps": round(2 * nbytes / el / 1e9, 1)}
    # the last grid gives each thread one 16-byte load and store
    for grid in (1024, 4096, 16384, 65536, nbytes // (16 * 256)):
        lib.fd23_device_copy(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), nbytes, grid, None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(lib.fd23_device_copy(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), nbytes,
                                            grid, None))
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / reps
        out["fd23_copy_grid%d_GBps" % grid] = round(2 * nbytes / el / 1e9, 1)
    del x, y
    # the passes' access pattern: arrays of 4096 x 16384 FP32 (256 MiB each,
    # above the 256 MB MALL) in one allocation, 64 x 16 tiles, timed with
    # HIP events
    nx, nz = 4096, 16384
    arrs = torch.zeros((11, nx * nz), dtype=torch.float32, device="cuda")
    ptr = [arrs[k].data_ptr() for k in range(11)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # (7, 2): the velocity pass's pattern, (8, 3): the stress pass's
    for nr, nw in ((1, 1), (3, 1), (7, 2), (8, 3)):
        rd = (ctypes.c_void_p * nr)(*ptr[:nr])
        wr = (ctypes.c_void_p * nw)(*ptr[8:8 + nw])
        _lib.check(lib.fd23_device_stream_tiles(rd, nr, wr, nw, nx, nz, None))
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            _lib.check(lib.fd23_device_stream_tiles(rd, nr, wr, nw, nx, nz, None))
        e1.record()
        torch.cuda.synchronize()
        el = e0.elapsed_time(e1) / 1e3 / reps
        out["tiles_r%d_w%d_GBps" % (nr, nw)] = round((nr + nw) * nx * nz * 4 / el / 1e9, 1)
    print(json.dumps(out), flush=True)


def bench_serve(nreq, max_active, rate=40.0, models=0, spare=None):
    """Replay the mixed request stream through the serving engine (Poisson
    arrivals at `rate` per second; models > 0: shot gathers on that many
    shared problems, see serving_stream, whose contexts keep `spare` idle
    shot slots (default 8) for gathers that arrive while they run)."""
    from fd23_amd import _lib
    from fd23_amd.serve import Server, replay, serving_stream

    stream = serving_stream(nreq, rate=rate, models=models)
    if spare is None:
        spare = 8 if models else 0
    with Server(max_active=max_active, chunk_steps=256, max_batch_shots=64, spare_shots=spare) as srv:
        recs, wall = replay(srv, stream)
        st = srv.stats()
    ok = [r for r in recs if r.get("ok")]
    lat = np.array([r["latency"] for r in ok])
    que = np.array([r["queue"] for r in ok])
    cells = sum(r["cells"] for r in ok)
    offered = sum(s.cells for _, s in stream) / max(stream[-1][0], 1e-9)
    return {
        "metric": "Gcell-updates/s",
        "value": round(cells / wall / 1e9, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic request stream (Poisson %g/s, seeded layered + GRF models)" % rate,
        "config": {"workload": "serving_mixed_sh_psv" if not models else "serving_shared_%d_models" % models,
                   "requests": nreq, "max_active": max_active, "spare_shots": spare, "chunk_steps": 256, "rate_per_s": rate},
        "requests_per_s": round(len(ok) / wall, 3),
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                      "max": round(float(lat.max()), 4)},
        "queue_s": {"p50": round(float(np.percentile(que, 50)), 4), "p99": round(float(np.percentile(que, 99)), 4)},
        "offered_gcell_per_s": round(offered / 1e9, 2),
        "service_s": {"p50": round(float(np.percentile(lat - que, 50)), 4),
                      "p99": round(float(np.percentile(lat - que, 99)), 4)},
        "wall_s": round(wall, 3),
        "completed": len(ok),
        "failed": len(recs) - len(ok),
        "finite": all(r["finite"] for r in ok),
        "engine": {k: st[k] for k in ("launches", "merged", "joined", "peak_active")},
        "mix": {"psv": sum(r["wave"] == _lib.WAVE_PSV for r in recs), "box": sum(r["box"] for r in recs),
                "f16": sum(r["storage"] == _lib.STORE_F16 for r in recs)},
    }


def bench_material(reps=5):
    """Material set-up of the 4096x2048 P-SV model: the host precompute
    plus its upload (fd23_psv_material + fd23_set_material) against the
    device precompute (fd23_set_model: three input arrays uploaded, one
    kernel); and for a 12-layer .mdl model, the host reader + host
    material + upload against fd23_set_model_file (polygons rasterised and
    the material computed on the device).  Median of reps, the context
    already created."""
    from fd23_amd import _lib, host, workloads as W
    from fd23_amd.core import Context

    wl = W.config_psv_precision(_lib.STORE_F32)
    vp, vs, den = W.layered_grf_model(wl.nx, wl.nz, wl.dx, wl.seed)
    t_host, t_dev = [], []
    with Context(_lib.WAVE_PSV, wl.nx, wl.nz, wl.npml) as ctx:
        for _ in range(reps):
            t0 = time.perf_counter()
            mat, _ = host.psv_material(vp, vs, den, wl.dx, wl.dt)
            ctx.set_material(mat)
            t_host.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            ctx.set_model(vp, vs, den, wl.dx, wl.dt)
            t_dev.append(time.perf_counter() - t0)
        same = all(np.array_equal(ctx.material(i), mat[i]) for i in range(5))
        # a .mdl model of 12 dipping layers: host reader + host material +
        # upload against the device rasteriser feeding the device material
        import tempfile

        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "layers.mdl")
            w, hgt = wl.nx * wl.dx, wl.nz * wl.dx
            lines = ["BACK 2000.0 1155.0 2100.0"]
            for k in range(1, 13):
                z0, z1 = k * hgt / 13, (k + 1) * hgt / 13 + 40.0
                lines += ["POLY %.1f %.1f %.1f" % (2000 + 300 * k, 1155 + 170 * k, 2100 + 40 * k),
                          "\t-10 %.1f" % z0, "\t%.1f %.1f" % (w + 10, z0 + 0.05 * w),
                          "\t%.1f %.1f" % (w + 10, z1 + 0.05 * w), "\t-10 %.1f" % z1]
            open(path, "w").write("\n".join(lines) + "\n")
            m_host, m_dev = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                mvp, mvs, mden = host.model_read(path, wl.nx, wl.nz, wl.dx)
                mmat, _ = host.psv_material(mvp, mvs, mden, wl.dx, wl.dt)
                ctx.set_material(mmat)
                m_host.append(time.perf_counter() - t0)
                t0 = time.perf_counter()
                ctx.set_model_file(path, wl.dx, wl.dt)
                m_dev.append(time.perf_counter() - t0)
            same = same and all(np.array_equal(ctx.material(i), mmat[i]) for i in range(5))
    h, d = float(np.median(t_host)) * 1e3, float(np.median(t_dev)) * 1e3
    mh, mdv = float(np.median(m_host)) * 1e3, float(np.median(m_dev)) * 1e3
    return {"metric": "material set-up time", "value": round(d, 3), "unit": "ms", "n_gpus": 1, "steps": None,
            "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded layered + GRF model)",
            "config": {"workload": "material_4096x2048", "nx": wl.nx, "nz": wl.nz, "reps": reps},
            "host_ms": round(h, 3), "device_ms": round(d, 3), "speedup": round(h / d, 2),
            "mdl_12_layers": {"host_ms": round(mh, 3), "device_ms": round(mdv, 3), "speedup": round(mh / mdv, 2)},
            "bitwise_equal": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--quick", action="store_true",
                    help="c4: the FP32 line only (no FP16/BF16, accuracy or batching figures)")
    ap.add_argument("--workload", default="c4",
                    choices=["psv2000", "batched", "single1024", "c4", "f16", "bf16", "sh400", "serve", "punch",
                             "material"])
    ap.add_argument("--requests", type=int, default=40, help="serve: requests in the stream")
    ap.add_argument("--max-active", type=int, default=8, help="serve: resident contexts")
    ap.add_argument("--rate", type=float, default=40.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--models", type=int, default=0, help="serve: shot gathers on this many shared problems")
    ap.add_argument("--spare-shots", type=int, default=None,
                    help="serve: idle shot slots a context keeps for joining requests (default 8 with --models)")
    ap.add_argument("--storage", choices=sorted(STORAGE), default=None)
    ap.add_argument("--accuracy", action="store_true",
                    help="also report seismogram error against an FP64 run of the same inputs")
    ap.add_argument("--accuracy-steps", type=int, default=0,
                    help="steps of the accuracy runs (default: 1000 for c4's line, else nt)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", type=int, default=None, help="use this tile shape instead of tuning")
    ap.add_argument("--chains", type=int, default=None,
                    help="with --tile: independent chains of passes over parts of the shots (1, 2 or 4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (shots and requests shard "
              "across GPUs with no communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args.requests, args.max_active, args.rate, args.models, args.spare_shots)),
              flush=True)
        return
    if args.workload == "material":
        print(json.dumps(bench_material()), flush=True)
        return

    from fd23_amd import _lib

    wl = make_workload(args.workload, args.storage).build()
    out = measure_line(wl, args.steps, args.warmup, autotune=not args.no_autotune, tile=args.tile,
                       chains=args.chains)
    headline = args.workload == "c4" and args.storage in (None, "f32") and not args.quick
    if headline:
        # FP16 / BF16 storage of the same workload, timed the same way
        variants = {}
        for st in ("f16", "bf16"):
            wl.storage = STORAGE[st]
            v = measure_line(wl, args.steps, args.warmup)
            variants[st] = {"value": v["value"], "unit": "Gcell/s", "ms_per_step": v["ms_per_step"],
                            "speedup_vs_f32": round(v["value"] / out["value"], 3), "tile": v["config"]["tile"],
                            "roofline_frac": v["roofline"]["frac"], "kernel_ms": v["roofline"]["kernel_ms"],
                            "kernel_name": v["roofline"]["kernel_name"], "finite": v["finite"]}
        wl.storage = STORAGE["f32"]
        acc = accuracy_vs_f64("c4", ("f32", "f16", "bf16"), args.accuracy_steps or 1000, built=wl)
        out["error_vs_fp64"] = acc["f32"]
        for st in ("f16", "bf16"):
            variants[st]["error_vs_fp64"] = acc[st]
        out["storage_variants"] = variants
        out["batching"] = bench_batching(max(args.steps, 64), max(args.warmup, 8))
    elif args.accuracy:
        sname = out["storage"]
        out["error_vs_fp64"] = accuracy_vs_f64(args.workload, (sname,), args.accuracy_steps or wl.nt)[sname]
    if not args.no_cpu and wl.wave == _lib.WAVE_PSV:
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)


def measure_line(wl, steps, warmup, autotune=True, tile=None, chains=None):
    """One workload timed over `steps` graph-replayed steps after `warmup`
    untimed ones, with the dominant pass's roofline."""
    from fd23_amd import _lib

    r = run_gpu(wl, steps, warmup, autotune=autotune, tile=tile, chains=chains)
    cells = float(wl.nshots) * wl.nx * wl.nz * steps
    gcell = cells / r["elapsed"] / 1e9
    ms = r["elapsed"] / steps * 1e3
    sb, vb = pass_bytes(wl)
    if r["fused"]:
        name, nbytes, kms, kname = "fused", fused_bytes(wl), r["stress_ms"], r["kernels"][0]
        step_bytes = nbytes
    elif r["block"]:
        # time_passes reports the time per step: one launch takes `block` steps
        k = r["block"]
        name, nbytes, kms, kname = "block", block_bytes(wl, k), r["stress_ms"] * k, r["kernels"][0]
        step_bytes = nbytes / k
    elif r["stress_ms"] >= r["vel_ms"]:
        name, nbytes, kms, kname = "stress", sb, r["stress_ms"], r["kernels"][0]
        step_bytes = sb + vb
    else:
        name, nbytes, kms, kname = "velocity", vb, r["vel_ms"], r["kernels"][1]
        step_bytes = sb + vb
    achieved = nbytes / (kms * 1e-3) / 1e9
    sname = {v: k for k, v in STORAGE.items()}[wl.storage]
    out = {
        "metric": "Gcell-updates/s",
        "value": round(gcell, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if sname in ("f32", "f16", "bf16") else "f64",
        "storage": sname,
        "data": "synthetic (seeded layered + GRF model, Ricker source)",
        "config": {"workload": workload_id(wl), "nx": wl.nx, "nz": wl.nz, "nshots": wl.nshots,
                   "npml": wl.npml, "dx": wl.dx, "dt": wl.dt, "nt": wl.nt, "tile": r["tile"],
                   "chains": r["chains"]},
        "shots_per_s": round(wl.nshots / (ms * 1e-3 * wl.nt), 4),
        "hbm_fraction_step": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "bytes_per_step": step_bytes,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(workload_id(wl), kname), "kernel_name": list(kname),
                     "kernel_ms": round(kms, 5), "algorithmic_bytes": nbytes,
                     "stress_ms": None if r["fused"] or r["block"] else round(r["stress_ms"], 5),
                     "velocity_ms": None if r["fused"] or r["block"] else round(r["vel_ms"], 5)},
        "finite": r["finite"],
    }
    if wl.punch_after is not None:
        # the same steps without the window, for the speed-up it gives
        before, after = wl.punch_before, wl.punch_after
        wl.punch_before = wl.punch_after = None
        r0 = run_gpu(wl, steps, warmup, autotune=False, tile=r["tile"])
        wl.punch_before, wl.punch_after = before, after
        out["first_arrival_window"] = {
            "before_s": before, "after_s": after,
            "active_fraction": round(wl.punch_active_fraction(warmup, steps), 4),
            "ms_per_step_without": round(r0["elapsed"] / steps * 1e3, 5),
            "speedup": round(r0["elapsed"] / r["elapsed"], 3),
            "note": "value counts the cells of the whole grid: an effective rate, not throughput"}
    return out


def bench_batching(steps, warmup, single_steps=512):
    """The 64-shot 1024x512 batch (one context, one launch per pass for all
    shots) against one 1024x512 shot, in shots/s over the configuration's
    4000 steps."""
    b = measure_line(make_workload("batched").build(), steps, warmup)
    one = measure_line(make_workload("single1024").build(), single_steps, warmup)
    return {"workload": "psv_1024x512_x64_fp32", "shots_per_s": b["shots_per_s"], "value": b["value"],
            "ms_per_step": b["ms_per_step"], "steps": steps, "tile": b["config"]["tile"],
            "chains": b["config"]["chains"], "roofline_frac": b["roofline"]["frac"],
            "kernel_ms": b["roofline"]["kernel_ms"], "kernel_name": b["roofline"]["kernel_name"],
            "single_shot": {"shots_per_s": one["shots_per_s"], "value": one["value"],
                            "ms_per_step": one["ms_per_step"], "steps": single_steps},
            "batch_over_single": round(b["shots_per_s"] / one["shots_per_s"], 3)}


if __name__ == "__main__":
    main()
