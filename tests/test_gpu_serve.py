"""Serving engine on the GPU: every request's seismograms are the same when
it runs alone and when it runs in a mix with others (admitted while others
are on the device, sharing a launch with a compatible request, overlapping
on other streams).

Tolerances, relative L2 per request against its alone run: FP32 and FP64
storage 1e-6 (the kernels are deterministic per shot, so in practice
bitwise); FP16 storage 1e-3 (the per-field power-of-two scales are shared
by the shots of one launch, so a merged launch may quantise with a
different scale).
"""
import time

import numpy as np
import pytest

from fd23_amd import _lib
from fd23_amd.core import Context
from fd23_amd.serve import Server, make_solve
from helpers import rel

pytestmark = pytest.mark.gpu

TOL = {_lib.STORE_F32: 1e-6, _lib.STORE_F64: 1e-6, _lib.STORE_F16: 1e-3, _lib.STORE_BF16: 1e-2}


def _requests():
    P, S = _lib.WAVE_PSV, _lib.WAVE_SH
    box = dict(x0=28, z0=4, bx=200, bz=100, incident=1, angle=20.0, f0=10.0)
    reqs = [
        make_solve(P, _lib.STORE_F32, 256, 128, 8.0, 400, 3, 11, tag="psv-a"),
        make_solve(P, _lib.STORE_F32, 256, 128, 8.0, 400, 2, 11, tag="psv-b"),   # same model as psv-a
        make_solve(S, _lib.STORE_F32, 320, 160, 6.0, 500, 2, 12, tag="sh"),
        make_solve(P, _lib.STORE_F16, 320, 192, 7.0, 450, 4, 13, tag="psv-f16"),
        make_solve(P, _lib.STORE_F32, 256, 128, 9.0, 300, 1, 14, box=box, tag="psv-box"),
        make_solve(S, _lib.STORE_F16, 256, 128, 5.0, 600, 3, 15, tag="sh-f16"),
        make_solve(P, _lib.STORE_F16, 256, 128, 6.0, 350, 1, 16, box=box, tag="psv-box-f16"),
    ]
    # psv-b shares psv-a's model (same seed, grid, dx) but has its own shots
    reqs[1].xs = reqs[1].xs + 7
    return reqs


def _alone(reqs):
    out = {}
    with Server(max_active=1, chunk_steps=64) as srv:
        for s in reqs:
            rid = srv.submit(s)
            assert srv.wait(rid, 120.0) == _lib.REQ_DONE
            out[s.tag] = srv.result(rid)[0]
    return out


def test_alone_equals_in_mix():
    reqs = _requests()
    print("case: serving alone vs mix, %d requests" % len(reqs), flush=True)
    alone = _alone(reqs)
    mix = {}
    with Server(max_active=4, chunk_steps=64, max_batch_shots=16) as srv:
        ids = []
        # a first wave at once (psv-a and psv-b queue together and merge),
        # the rest while the first ones are on the device
        for s in reqs[:3]:
            ids.append((s, srv.submit(s)))
        time.sleep(0.02)
        for s in reqs[3:]:
            ids.append((s, srv.submit(s)))
            time.sleep(0.005)
        for s, rid in ids:
            assert srv.wait(rid, 180.0) == _lib.REQ_DONE, s.tag
            seis, lat, que = srv.result(rid)
            mix[s.tag] = seis
            assert lat >= que >= 0.0
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_active"] >= 2
    for s in reqs:
        a, m = alone[s.tag], mix[s.tag]
        assert a.shape == m.shape == (s.nshots, s.ncomp, s.nt, s.rx.size)
        assert np.isfinite(m).all() and np.abs(a).max() > 0
        eff = _lib.STORE_F32 if (s.box and s.storage == _lib.STORE_F16) else s.storage
        e = rel(m, a)
        print(s.tag, "alone vs mix", e, flush=True)
        assert e <= TOL[eff], s.tag


def test_merged_requests_share_a_launch():
    reqs = _requests()
    blocker, a, b = reqs[5], reqs[0], reqs[1]
    with Server(max_active=1, chunk_steps=64, max_batch_shots=16) as srv:
        # the only slot is busy with the blocker while a and b queue, so the
        # next admission round finds both and merges them into one launch
        ids = [srv.submit(blocker), srv.submit(a), srv.submit(b)]
        for rid in ids:
            assert srv.wait(rid, 120.0) == _lib.REQ_DONE
            srv.result(rid)
        st = srv.stats()
    assert st["launches"] == 2 and st["merged"] == 2, st


def test_a_request_joins_a_running_context():
    """Requests on one problem with different numbers of steps share a
    context; a request that arrives after the shorter one has finished takes
    its idle shot slot while the longer one still runs (the only stream slot
    is busy, so joining is its one way in).  Each result equals its alone
    run."""
    P, F32 = _lib.WAVE_PSV, _lib.STORE_F32
    blocker = _requests()[5]
    long_ = make_solve(P, F32, 256, 128, 8.0, 3000, 2, 31, tag="long")
    short = make_solve(P, F32, 256, 128, 8.0, 200, 1, 31, tag="short")
    late = make_solve(P, F32, 256, 128, 8.0, 300, 1, 31, tag="late")
    short.xs = short.xs + 11
    late.xs = late.xs + 23
    alone = _alone([long_, short, late])
    with Server(max_active=1, chunk_steps=64, max_batch_shots=16) as srv:
        ids = {"blocker": srv.submit(blocker), "long": srv.submit(long_), "short": srv.submit(short)}
        assert srv.wait(ids["short"], 120.0) == _lib.REQ_DONE
        ids["late"] = srv.submit(late)
        out = {}
        for tag, rid in ids.items():
            assert srv.wait(rid, 180.0) == _lib.REQ_DONE, tag
            out[tag] = srv.result(rid)[0]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["launches"] == 2 and st["merged"] == 2 and st["joined"] == 1, st
    for s in (long_, short, late):
        assert out[s.tag].shape == alone[s.tag].shape
        np.testing.assert_array_equal(out[s.tag], alone[s.tag])


def test_an_arrival_joins_spare_slots():
    """A context for a shared problem keeps idle shot slots (spare_shots);
    a gather that arrives while it runs starts in them at the next chunk
    boundary instead of waiting for a stream slot.  Both results equal their
    alone runs."""
    P, F32 = _lib.WAVE_PSV, _lib.STORE_F32
    first = make_solve(P, F32, 256, 128, 8.0, 6000, 2, 33, tag="first")
    later = make_solve(P, F32, 256, 128, 8.0, 700, 3, 33, tag="later")
    later.xs = later.xs + 5
    alone = _alone([first, later])
    with Server(max_active=1, chunk_steps=64, max_batch_shots=16, spare_shots=4) as srv:
        a = srv.submit(first)
        t0 = time.time()
        while srv.wait(a, 0.0) == _lib.REQ_QUEUED:  # until the first request is admitted
            assert time.time() - t0 < 60.0
            time.sleep(0.001)
        b = srv.submit(later)
        out = {}
        for tag, rid in (("first", a), ("later", b)):
            assert srv.wait(rid, 180.0) == _lib.REQ_DONE, tag
            out[tag] = srv.result(rid)[0]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["launches"] == 1 and st["joined"] == 1, st
    for s in (first, later):
        np.testing.assert_array_equal(out[s.tag], alone[s.tag])


def test_server_matches_direct_context():
    s = _requests()[2]
    with Server(max_active=1) as srv:
        rid = srv.submit(s)
        assert srv.wait(rid, 120.0) == _lib.REQ_DONE
        seis = srv.result(rid)[0]
    ctx = Context(s.wave, s.nx, s.nz, s.npml, nshots=s.nshots, storage=s.storage, max_frames=s.nt,
                  max_rec=s.rx.size)
    ctx.set_material(s.material)
    ctx.set_cpml(s.cpml)
    for k in range(s.nshots):
        ctx.set_source(k, s.src_kind, int(s.xs[k]), int(s.zs[k]), s.stf, s.M)
        ctx.set_receivers(k, 0, s.rx, s.rz)
    ctx.run(0, s.nt)
    for k in range(s.nshots):
        np.testing.assert_array_equal(seis[k, 0], ctx.seismograms(k, 0, s.nt))
    ctx.close()


def test_admission_by_budget():
    reqs = _requests()
    small = reqs[0]
    from fd23_amd.core import desc_device_bytes

    need = desc_device_bytes(small.wave, small.nx, small.nz, small.npml, small.nshots, small.storage,
                             max_frames=small.nt, max_rec=small.rx.size)
    # room for about one such request: requests run one after another
    with Server(hbm_budget=int(need * 1.25) + (2 << 20), max_active=4, chunk_steps=64) as srv:
        ids = [srv.submit(reqs[0]), srv.submit(reqs[2])]
        big = make_solve(_lib.WAVE_PSV, _lib.STORE_F32, 1024, 512, 5.0, 100, 8, 99, tag="big")
        bid = srv.submit(big)
        for rid in ids:
            assert srv.wait(rid, 120.0) == _lib.REQ_DONE
            srv.result(rid)
        assert srv.wait(bid, 30.0) == _lib.REQ_FAILED
        with pytest.raises(_lib.Fd23Error):
            srv.result(bid)
        st = srv.stats()
    assert st["peak_active"] == 1 and st["failed"] == 1


def test_requests_with_model_equal_requests_with_material():
    """A request that carries vp, vs, den (the engine computes the material
    on the device at admission) returns the same seismograms, bitwise, as
    the same request carrying the host-computed material."""
    P, S = _lib.WAVE_PSV, _lib.WAVE_SH
    print("case: serving requests with the model", flush=True)
    # (FP64 storage: make_solve's material is computed in float, the
    # device's from the model in double; test_gpu_material covers FP64)
    specs = [(P, _lib.STORE_F32, 256, 128, 8.0, 300, 2, 31), (S, _lib.STORE_F32, 192, 96, 6.0, 300, 1, 32),
             (P, _lib.STORE_F16, 256, 128, 7.0, 300, 2, 34)]
    with Server(max_active=4, chunk_steps=64) as srv:
        for spec in specs:
            a = make_solve(*spec, tag="material")
            b = make_solve(*spec, tag="model", device_material=True)
            assert b.material is None and b.model is not None
            got = []
            for s in (a, b):
                rid = srv.submit(s)
                assert srv.wait(rid, 120.0) == _lib.REQ_DONE
                got.append(srv.result(rid)[0])
            assert np.abs(got[0]).max() > 0
            assert np.array_equal(got[0], got[1]), spec


# ---------------------------------------------------------------------------
# a slice of the serving configuration's stream, served and pinned to the
# oracle
# ---------------------------------------------------------------------------
C5_NT_CAP, C5_SHOT_CAP = 1500, 3


def _c5_slice(n=12):
    """The first n requests of the serving stream (seed 5005: every grid
    size of the mix, plane-wave boxes, FP16 targets), with the step count
    capped at 1500 and the shots at 3 to bound the test; every request
    keeps at least 1000 steps."""
    from fd23_amd.serve import serving_stream

    out = []
    for at, s in serving_stream(n):
        s.nt = min(s.nt, C5_NT_CAP)
        s.xs, s.zs = s.xs[:C5_SHOT_CAP], s.zs[:C5_SHOT_CAP]
        out.append((at, s))
    return out


def _oracle_request(s, shot, dtype, fmt=None):
    """Seismograms [comp][frame][rec] of one shot of a served request on the
    oracle: point source, plane-wave box when the request has one, FP32 or
    FP64 arithmetic; fmt: FP16/BF16 storage emulated after every pass."""
    import oracle as O
    from helpers import unpack_cpml

    cp = unpack_cpml(s.cpml, s.wave, s.nx, s.nz, dtype)
    xs, zs = int(s.xs[shot]), int(s.zs[shot])
    stf, M = s.stf.astype(dtype), s.M.astype(dtype)
    if s.wave == _lib.WAVE_PSV:
        sim = O.PSV(s.nx, s.nz, s.npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), [m.copy() for m in s.material])),
                    cp, s.topabsorb, dtype)
        kind, comps = 0, ("U", "W")
        after_s = lambda: [sim.Txx, sim.Tzz, sim.Txz] + [sim.psi[n] for n in ("U_x", "W_x", "U_z", "W_z")]
        after_v = lambda: [sim.U, sim.W] + [sim.psi[n] for n in ("Txx_x", "Txz_x", "Txz_z", "Tzz_z")]
    else:
        sim = O.SH(s.nx, s.nz, s.npml, dict(zip(("BV", "MUX", "MUZ"), s.material)), cp, s.topabsorb, dtype)
        kind, comps = s.src_kind, ("V",)
        after_s = lambda: [sim.Tx, sim.Tz, sim.psi["V_x"], sim.psi["V_z"]]
        after_v = lambda: [sim.V, sim.psi["Tx_x"], sim.psi["Tz_z"]]
    box = None
    if s.box:
        b = s.box
        box = O.PlaneWaveBox(sim, b["x0"], b["z0"], b["bx"], b["bz"], b["incident"], b["angle"], b["f0"], b["t0"],
                             b["amp"], b["vp"], b["vs"], b["rho"], s.dx, s.dt)
    seis = [[] for _ in comps]
    for it in range(s.nt):
        if box:
            box.pre_stress(it)
        sim.stress(it)
        sim.inject(kind, it, stf, M, xs, zs)
        if fmt:
            for a in after_s():
                O.round_storage(a, fmt)
        if box:
            box.post_stress(it)
        sim.velocity(it)
        if fmt:
            for a in after_v():
                O.round_storage(a, fmt)
        for c, nm in enumerate(comps):
            seis[c].append(sim.record(nm, s.rx, s.rz))
        if box:
            box.post_velocity(it)
    return np.array(seis)


def test_serving_stream_slice_against_the_oracle():
    """Twelve requests of the serving stream submitted on their arrival
    schedule (compressed) to one engine: FP32 requests up to 512x256 match
    the FP32 oracle within 1e-5 (every shot); requests with a plane-wave box
    (FP16 ones run at FP32 storage) match the FP64 oracle within 1.5x the
    FP32 oracle's own error against it (shot 0); FP16 requests up to 512x256 match the
    FP64 oracle within 1.3x the FP16 storage floor emulated on it (shot 0);
    larger requests equal their alone runs."""
    from fd23_amd.serve import Server

    stream = _c5_slice()
    sizes = {(s.nx, s.nz) for _, s in stream}
    assert len(sizes) == 4 and any(s.box for _, s in stream)
    assert any(s.storage == _lib.STORE_F16 for _, s in stream) and min(s.nt for _, s in stream) >= 1000
    print("case: serving stream slice, %d requests" % len(stream), flush=True)
    got = {}
    with Server(max_active=4, chunk_steps=256) as srv:
        ids = []
        t0 = time.perf_counter()
        for at, s in stream:
            delay = at / 10.0 - (time.perf_counter() - t0)  # arrivals at 10x the stream's rate
            if delay > 0:
                time.sleep(delay)
            ids.append((s, srv.submit(s)))
        for s, rid in ids:
            assert srv.wait(rid, 300.0) == _lib.REQ_DONE, s.tag
            got[s.tag] = srv.result(rid)[0]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["completed"] == len(stream) and st["peak_active"] >= 2
    big = [s for _, s in stream if s.nx * s.nz > 512 * 256]
    alone = _alone(big)
    for _, s in stream:
        g = got[s.tag]
        assert np.isfinite(g).all() and np.abs(g).max() > 0, s.tag
        eff = _lib.STORE_F32 if (s.box and s.storage == _lib.STORE_F16) else s.storage
        if s.nx * s.nz > 512 * 256:
            e = rel(g, alone[s.tag])
            print(s.tag, s.nx, s.nz, "served vs alone", e, flush=True)
            assert e <= TOL[eff], s.tag
        elif eff == _lib.STORE_F32 and s.box:
            # outside the box the global grid holds the scattered field, a
            # small difference of amplitude-1 incident values: FP32 rounding
            # of the incident wave sets the error floor there, so shot 0 is
            # measured against the FP64 oracle beside the FP32 oracle's own
            # error (the drive is numpy in the oracle: shot 0 only)
            ref = _oracle_request(s, 0, np.float64)
            e, floor = rel(g[0].astype(np.float64), ref), rel(_oracle_request(s, 0, np.float32), ref)
            print(s.tag, s.nx, s.nz, "box: served vs FP64 oracle", e, "FP32 oracle vs FP64", floor, flush=True)
            assert e <= max(1e-5, 1.5 * floor), s.tag
        elif eff == _lib.STORE_F32:
            errs = [rel(g[k], _oracle_request(s, k, np.float32)) for k in range(s.nshots)]
            print(s.tag, s.nx, s.nz, "served vs FP32 oracle", errs, flush=True)
            assert max(errs) < 1e-5, s.tag
        else:
            ref = _oracle_request(s, 0, np.float64)
            floor = rel(_oracle_request(s, 0, np.float64, fmt="f16"), ref)
            e = rel(g[0].astype(np.float64), ref)
            print(s.tag, s.nx, s.nz, "served FP16 vs FP64 oracle", e, "floor", floor, flush=True)
            assert e <= 1.3 * floor, s.tag
