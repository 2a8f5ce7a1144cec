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
