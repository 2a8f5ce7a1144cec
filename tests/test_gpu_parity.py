"""HIP path vs the oracle and vs the reference's own checkpoint fixtures.

All calls go through libfd23.so's C ABI.  Tolerances: 1e-5 relative L2 on
seismograms and fields for FP32 (the stated parity bar), 1e-6 on
the reference's step-0/1 checkpoints, 1e-10 for FP64 storage vs the FP64
oracle.
"""
import os

import numpy as np
import pytest

import oracle as O
from fd23_amd import _lib, host
from fd23_amd.core import Context
from helpers import Problem, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref")


def _print(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------------------
# reference checkpoints: PSV/clean_nobox/TEST/tt (96x96, npml 32, top C-PML)
# ---------------------------------------------------------------------------
def _tt_context():
    n, npml, h, dt = 96, 32, 0.3, 0.01
    vp, vs, den = host.model_read(os.path.join(GOLD, "psv_tt", "homo_6.8.mdl"), n, n, h)
    mat, stab = host.psv_material(vp, vs, den, h, dt)
    packed = host.cpml_pack(_lib.WAVE_PSV, n, n, npml, 0.005, 1e-11, 30.0, 2.0)
    ctx = Context(_lib.WAVE_PSV, n, n, npml, topabsorb=True, max_frames=1000, max_rec=2)
    ctx.set_material(mat)
    ctx.set_cpml(packed)
    stf = host.gaussian(-30)
    M = host.psv_moment(_lib.SRC_LINE, 0, 20, 90, 45, h, dt)
    ctx.set_source(0, _lib.SRC_LINE, 33, 33, stf, M)
    return ctx


def _state(name):
    return np.fromfile(os.path.join(GOLD, "psv_tt", name), np.float32)


def test_reference_checkpoints_step0_step1():
    _print("case: tt checkpoints step 0/1")
    with _tt_context() as ctx:
        for it, name in ((0, "state000000"), (1, "state000001")):
            ctx.step_stress()
            ctx.inject(it)
            ctx.step_velocity()
            got = np.frombuffer(ctx.get_state(), np.float32)
            err = rel(got, _state(name))
            _print(name, err)
            assert err < 1e-6


def test_reference_checkpoint_500_steps_and_restart():
    _print("case: tt 500 steps + restart")
    with _tt_context() as ctx:
        ctx.run(0, 500, 1)
        got = np.frombuffer(ctx.get_state(), np.float32)
        err = rel(got, _state("state000499"))
        _print("state000499", err)
        assert err < 1e-5
    with _tt_context() as ctx:
        ctx.set_state(_state("state000499").tobytes())
        ctx.run(500, 500, 1)
        got = np.frombuffer(ctx.get_state(), np.float32)
        err = rel(got, _state("state000999"))
        _print("state000999", err)
        assert err < 1e-5


# ---------------------------------------------------------------------------
# oracle parity at non-multiple sizes, all edge branches
# ---------------------------------------------------------------------------
PSV_CASES = [
    # nx, nz, npml, top, kind, explosion
    (150, 90, 12, False, _lib.SRC_LINE, False),
    (131, 77, 10, True, _lib.SRC_LINE, False),
    (200, 64, 16, False, _lib.SRC_P, False),
    (97, 120, 8, True, _lib.SRC_LINE, True),
    (24, 16, 3, False, _lib.SRC_LINE, False),  # the smallest grid side (16), smaller than one tile
    (65, 33, 5, True, _lib.SRC_LINE, False),   # one column and one row past a tile
]


@pytest.mark.parametrize("nx,nz,npml,top,kind,expl", PSV_CASES)
def test_psv_vs_oracle(nx, nz, npml, top, kind, expl):
    _print("case: psv", nx, nz, npml, top, kind, expl)
    # receivers below the frozen top C-PML rows when the top absorbs
    p = Problem(_lib.WAVE_PSV, nx, nz, npml, 10.0, 7e-4, 400, seed=nx, topabsorb=top, kind=kind,
                explosion=expl, sources=[(nx // 3, 6)], rec_z=npml + 3 if top else 3)
    with p.context() as ctx:
        ctx.run(0, p.nt, 1)
        gu, gw = ctx.seismograms(0, 0, p.nt), ctx.seismograms(0, 1, p.nt)
        fields = {n: ctx.field(n) for n in ("U", "W", "Txx", "Txz", "Tzz")}
        st = np.frombuffer(ctx.get_state(), np.float32)
    sim, (ou, ow) = p.oracle()
    eu, ew = rel(gu, ou), rel(gw, ow)
    _print("seis", eu, ew)
    assert eu < 1e-5 and ew < 1e-5
    for n, f in fields.items():
        e = rel(f, getattr(sim, n))
        _print(n, e)
        assert e < 1e-5
    ost = np.frombuffer(sim.state_bytes(), np.float32)
    assert rel(st, ost) < 1e-5


SH_CASES = [
    (150, 90, 12, False, _lib.SRC_LINE),
    (131, 77, 10, True, _lib.SRC_P),
    (120, 100, 9, True, _lib.SRC_FORCE),
    (17, 16, 3, False, _lib.SRC_FORCE),  # the smallest grid side (16), smaller than one tile
]


@pytest.mark.parametrize("nx,nz,npml,top,kind", SH_CASES)
def test_sh_vs_oracle(nx, nz, npml, top, kind):
    _print("case: sh", nx, nz, npml, top, kind)
    p = Problem(_lib.WAVE_SH, nx, nz, npml, 10.0, 7e-4, 400, seed=nx, topabsorb=top, kind=kind,
                sources=[(nx // 2, 8)], rec_z=npml + 2 if top else 2)
    with p.context() as ctx:
        ctx.run(0, p.nt, 1)
        gv = ctx.seismograms(0, 0, p.nt)
        fields = {n: ctx.field(n) for n in ("V", "Tx", "Tz")}
    sim, (ov,) = p.oracle()
    e = rel(gv, ov)
    _print("seis", e)
    assert e < 1e-5
    for n, f in fields.items():
        e = rel(f, getattr(sim, n))
        _print(n, e)
        assert e < 1e-5


def test_psv_f64_storage_vs_f64_oracle():
    _print("case: psv f64")
    p = Problem(_lib.WAVE_PSV, 140, 80, 10, 10.0, 7e-4, 150, seed=3, f64=True, sources=[(70, 6)])
    with p.context() as ctx:
        ctx.run(0, p.nt, 1)
        gu, gw = ctx.seismograms(0, 0, p.nt), ctx.seismograms(0, 1, p.nt)
    sim, (ou, ow) = p.oracle()
    _print(rel(gu, ou), rel(gw, ow))
    assert rel(gu, ou) < 1e-10 and rel(gw, ow) < 1e-10


# ---------------------------------------------------------------------------
# internal consistency: graph loop == staged calls, tiles agree, batch == solo
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("wave,kind", [(_lib.WAVE_PSV, _lib.SRC_LINE), (_lib.WAVE_PSV, _lib.SRC_P),
                                       (_lib.WAVE_SH, _lib.SRC_FORCE), (_lib.WAVE_SH, _lib.SRC_P)])
def test_graph_run_equals_staged_calls(wave, kind):
    """fd23_run (graph loop, injection and gather fused into the passes) is
    bitwise the staged reference order stress / inject / velocity / record."""
    _print("case: graph vs staged", wave, kind)
    # 140 steps: one full 128-step graph chunk and a graph of the rest
    p = Problem(wave, 130, 70, 10, 10.0, 7e-4, 140, seed=5, kind=kind,
                sources=[(60, 6), (61, 30)], rec_z=4)
    with p.context() as a, p.context() as b:
        a.run(0, p.nt, 2)
        for it in range(p.nt):
            b.step_stress()
            b.inject(it)
            b.step_velocity()
            b.record(it, 2)
        fr = (p.nt + 1) // 2
        for s in range(2):
            assert a.get_state(s) == b.get_state(s)
            assert np.array_equal(a.seismograms(s, 0, fr), b.seismograms(s, 0, fr))


def test_all_tiles_bitwise_equal():
    _print("case: tiles")
    p = Problem(_lib.WAVE_PSV, 150, 90, 12, 10.0, 7e-4, 40, seed=9, sources=[(40, 6), (100, 9)])
    ref = None
    lib = _lib.load()
    for t in range(lib.fd23_num_tiles()):
        if lib.fd23_tile_kind(t) == _lib.TILE_BLOCK:  # SH only
            continue
        with p.context() as ctx:
            ctx.tile = t
            ctx.run(0, p.nt, 1)
            st = ctx.get_state(0) + ctx.get_state(1)
            st += ctx.seismograms(1, 0, p.nt).tobytes() + ctx.seismograms(0, 1, p.nt).tobytes()
        if ref is None:
            ref = st
        _print("tile", t, st == ref)
        assert st == ref


def test_chains_bitwise_equal_one_chain():
    """Chunks captured as 2 or 4 chains of passes over parts of the shot
    groups give states and seismograms bitwise equal to one chain, on
    unsplit, split and several-shots-per-workgroup shapes, over a full
    128-step chunk and a remainder, with 5 shots (parts of unequal size,
    an empty one where 4 chains share 3 groups)."""
    lib = _lib.load()
    srcs = [(30, 6), (60, 9), (90, 4), (120, 12), (45, 20)]
    p = Problem(_lib.WAVE_PSV, 150, 90, 12, 10.0, 7e-4, 150, seed=17, kind=_lib.SRC_P, sources=srcs, rec_z=3)
    # RZ 2 and 4 unsplit, 4 shots per workgroup (two groups: 4 and 1 shots),
    # RZ 2 split, 2 shots per workgroup split
    for t in (0, 1, 4, 14, 16):
        assert lib.fd23_tile_kind(t) == _lib.TILE_TILED
        out = []
        for n in (1, 2, 4):
            with p.context() as ctx:
                ctx.tile = t
                ctx.chains = n
                ctx.run(0, p.nt, 1)
                groups = -(-len(srcs) // ctx.tile_shape()[1])
                assert ctx.chains == (n if groups >= n else 1)
                ctx.sync()
                out.append((b"".join(ctx.get_state(s) for s in range(len(srcs))),
                            [ctx.seismograms(s, c, p.nt) for s in range(len(srcs)) for c in range(2)]))
        for k in (1, 2):
            _print("chains tile", t, (1, 2, 4)[k], out[0][0] == out[k][0])
            assert out[0][0] == out[k][0]
            for a_, b_ in zip(out[0][1], out[k][1]):
                assert np.array_equal(a_, b_)


def test_chains_apply_only_where_they_can():
    """fd23_set_chains takes 0, 1, 2 or 4; a context runs one chain where
    chains cannot apply: a fused shape, 16-bit storage (scales shared by
    all shots), a single shot group."""
    lib = _lib.load()
    srcs = [(30, 6), (60, 9), (90, 4), (45, 20)]
    p = Problem(_lib.WAVE_PSV, 128, 64, 8, 10.0, 7e-4, 20, seed=3, sources=srcs)
    with p.context() as ctx:
        for bad in (-1, 3, 5):
            assert lib.fd23_set_chains(ctx.h, bad) == _lib.EINVAL
        ctx.tile = 0
        ctx.chains = 4
        assert ctx.chains == 4
        ctx.chains = 2
        assert ctx.chains == 2
        ctx.tile = _fused_tiles()[0]
        assert ctx.chains == 1
        ctx.tile = 5  # 8 shots per workgroup: one group of 4 shots
        assert ctx.chains == 1
    with p.context(storage=_lib.STORE_F16) as ctx:
        ctx.tile = 0
        ctx.chains = 2
        assert ctx.chains == 1
    q = Problem(_lib.WAVE_PSV, 128, 64, 8, 10.0, 7e-4, 20, seed=3, sources=srcs[:1])
    with q.context() as ctx:
        ctx.chains = 2
        assert ctx.chains == 1


def _fused_tiles():
    """Fused P-SV shapes: the strip-marching step (segments of 32/64/128 rows
    in blocks of 4, 64 rows in 2) and the tile-parallel step (segments of
    16 and 32 rows, block rows 0)."""
    lib = _lib.load()
    return [t for t in range(lib.fd23_num_tiles()) if lib.fd23_tile_kind(t) == _lib.TILE_FUSED]


@pytest.mark.parametrize("nx,nz,npml,top,kind,storage", [
    (173, 150, 12, False, _lib.SRC_LINE, _lib.STORE_F32),
    (130, 97, 10, True, _lib.SRC_P, _lib.STORE_F32),
    (121, 140, 9, False, _lib.SRC_LINE, _lib.STORE_F64),
])
def test_fused_step_bitwise_equals_two_pass(nx, nz, npml, top, kind, storage):
    """The fused step runs the tiled kernels' per-cell code: states and
    seismograms must be identical, across buffer swaps (odd step counts),
    split runs, and staged steps taken after fused ones; receivers sit on
    strip edges (strips own 56 columns)."""
    _print("case: fused", nx, nz, npml, top, kind, storage)
    srcs = [(nx // 2, 7), (5, 3), (nx - 6, nz // 2)]
    p = Problem(_lib.WAVE_PSV, nx, nz, npml, 10.0, 7e-4, 171, seed=21, topabsorb=top, kind=kind,
                sources=srcs, rec_z=npml + 3 if top else 2, f64=storage == _lib.STORE_F64)
    p.rx = np.concatenate([p.rx, np.array([0, nx - 1, 55, 56, 111, 112], np.int32)])
    p.rz = np.concatenate([p.rz, np.array([nz - 1, 0, nz // 2, nz // 2, 60, 61], np.int32)])

    def run(tile, staged_tail):
        with p.context(storage=storage) as ctx:
            ctx.tile = tile
            ctx.run(0, 133, 1)   # one 128-step chunk and 5 steps
            ctx.run(133, 28, 1)
            for it in range(161, 161 + staged_tail):
                ctx.step_stress()
                ctx.inject(it)
                ctx.step_velocity()
                ctx.record(it)
            if staged_tail < 10:
                ctx.run(161 + staged_tail, 10 - staged_tail, 1)
            ctx.sync()
            st = b"".join(ctx.get_state(s) for s in range(len(srcs)))
            seis = [ctx.seismograms(s, c, p.nt) for s in range(len(srcs)) for c in range(2)]
        return st, seis

    ref_st, ref_seis = run(0, 3)
    assert any(np.abs(x).max() > 0 for x in ref_seis)
    for t in _fused_tiles():
        st, seis = run(t, 3)
        _print("fused tile", t, st == ref_st)
        assert st == ref_st
        for a, b in zip(seis, ref_seis):
            assert np.array_equal(a, b)


def _block_tiles():
    lib = _lib.load()
    return [t for t in range(lib.fd23_num_tiles()) if lib.fd23_tile_kind(t) == _lib.TILE_BLOCK]


@pytest.mark.parametrize("nx,nz,npml,top,kind,storage", [
    (173, 150, 12, False, _lib.SRC_FORCE, _lib.STORE_F32),
    (130, 97, 10, True, _lib.SRC_P, _lib.STORE_F32),
    (121, 60, 9, True, _lib.SRC_LINE, _lib.STORE_F64),
    (17, 16, 3, False, _lib.SRC_FORCE, _lib.STORE_F32),  # smaller than one tile
])
def test_sh_block_step_bitwise_equals_two_pass(nx, nz, npml, top, kind, storage):
    """The SH block step (up to K steps per launch over tiles whose halos
    the launch recomputes, two buffer sets) runs the two-pass kernels'
    per-cell code: states and seismograms must be identical across chunks
    that end on either buffer set (replayed from cached graphs too),
    launches of fewer steps than K, staged steps between block runs, and
    sources and receivers on tile edges, grid edges and the surface."""
    _print("case: sh block", nx, nz, npml, top, kind, storage)
    srcs = [(nx // 2, 7), (3, 2), (nx - 5, nz // 2)]
    p = Problem(_lib.WAVE_SH, nx, nz, npml, 10.0, 7e-4, 186, seed=31, topabsorb=top, kind=kind,
                sources=srcs, rec_z=npml + 2 if top else 1, f64=storage == _lib.STORE_F64)
    # tile edges of the block shapes: 50 or 36 columns, 10/18/20/34/36 rows
    ex = np.array([0, nx - 1, 49, 50, 35, 36, 99, 100, 3, 5, 17, 18], np.int32) % nx
    ez = np.array([nz - 1, 0, 9, 10, 17, 18, 19, 20, 33, 34, 35, 36], np.int32) % nz
    p.rx = np.concatenate([p.rx, ex])
    p.rz = np.concatenate([p.rz, ez])

    def run(tile):
        with p.context(storage=storage) as ctx:
            ctx.tile = tile
            ctx.run(0, 133, 1)   # one 128-step chunk and 5 steps
            ctx.run(133, 28, 1)
            for it in range(161, 164):
                ctx.step_stress()
                ctx.inject(it)
                ctx.step_velocity()
                ctx.record(it)
            ctx.run(164, 7, 1)
            for it0 in (171, 176, 181):  # the same 5-step chunk three times
                ctx.run(it0, 5, 1)
            ctx.sync()
            st = b"".join(ctx.get_state(s) for s in range(len(srcs)))
            seis = [ctx.seismograms(s, 0, p.nt) for s in range(len(srcs))]
        return st, seis

    ref_st, ref_seis = run(0)
    assert any(np.abs(x).max() > 0 for x in ref_seis)
    for t in _block_tiles():
        st, seis = run(t)
        _print("block tile", t, st == ref_st)
        assert st == ref_st
        for a, b in zip(seis, ref_seis):
            assert np.array_equal(a, b)


def test_sh_block_step_restart_from_saved_state():
    """A block-step run stopped after a chunk that leaves the state in the
    second buffer set, saved, and continued in a new context from that
    state equals the uninterrupted run bitwise (state and the seismograms
    after the restart)."""
    _print("case: sh block restart")
    p = Problem(_lib.WAVE_SH, 150, 90, 12, 10.0, 7e-4, 80, seed=41, kind=_lib.SRC_FORCE,
                sources=[(75, 8), (30, 20)], rec_z=2)
    t = _block_tiles()[0]
    with p.context() as a:
        a.tile = t
        a.run(0, 37, 1)
        saved = [a.get_state(s) for s in range(2)]
        a.run(37, 43, 1)
        ref = [a.get_state(s) for s in range(2)]
        ref_seis = [a.seismograms(s, 0, p.nt)[37:] for s in range(2)]
    with p.context() as b:
        b.tile = t
        for s in range(2):
            b.set_state(saved[s], s)
        b.run(37, 43, 1)
        assert [b.get_state(s) for s in range(2)] == ref
        for s in range(2):
            assert np.array_equal(b.seismograms(s, 0, p.nt)[37:], ref_seis[s])


def test_sh_block_step_vs_oracle():
    """The block step on its own against the oracle (FP32, 1e-5)."""
    _print("case: sh block vs oracle")
    p = Problem(_lib.WAVE_SH, 150, 90, 12, 10.0, 7e-4, 400, seed=150, kind=_lib.SRC_FORCE,
                sources=[(75, 8)], rec_z=2)
    sim, (ov,) = p.oracle()
    for t in _block_tiles():
        with p.context() as ctx:
            ctx.tile = t
            ctx.run(0, p.nt, 1)
            gv = ctx.seismograms(0, 0, p.nt)
            fields = {n: ctx.field(n) for n in ("V", "Tx", "Tz")}
        e = rel(gv, ov)
        _print("tile", t, "seis", e)
        assert e < 1e-5
        for n, f in fields.items():
            assert rel(f, getattr(sim, n)) < 1e-5


def test_block_tiles_refused_where_they_do_not_apply():
    """SH block shapes on P-SV or FP16 contexts, and the tile-parallel
    fused shapes (block rows 0) on FP16 contexts, are refused with an
    error naming the path; the staged entry points and tile 0 still run."""
    p = Problem(_lib.WAVE_PSV, 64, 32, 4, 10.0, 7e-4, 4, seed=1)
    q = Problem(_lib.WAVE_SH, 64, 32, 4, 10.0, 7e-4, 4, seed=1)
    t = _block_tiles()[0]
    with p.context() as ctx, pytest.raises(Exception, match="block step"):
        ctx.tile = t
    with q.context(storage=_lib.STORE_F16) as ctx, pytest.raises(Exception, match="block step"):
        ctx.tile = t
    with p.context(storage=_lib.STORE_F16) as ctx:
        ftile = [f for f in _fused_tiles() if ctx.tile_shape(f)[3] == 0]
        assert ftile
        with pytest.raises(Exception, match="tile-parallel fused step"):
            ctx.tile = ftile[0]
        ctx.run(0, p.nt, 1)
        assert np.isfinite(ctx.seismograms(0, 0, p.nt)).all()


def test_fused_step_f16_close_to_two_pass():
    """FP16 storage: the fused step uses the unrounded new stresses in the
    velocity update, so it differs from the two-pass path by FP16 rounding
    noise (the storage error of either path against FP64 is ~4e-3 here)."""
    _print("case: fused f16")
    p = Problem(_lib.WAVE_PSV, 200, 160, 16, 10.0, 7e-4, 300, seed=23, sources=[(100, 8), (60, 12)], rec_z=5)
    out = {}
    for t in (0, 11):
        with p.context(storage=_lib.STORE_F16) as ctx:
            ctx.tile = t
            ctx.run(0, p.nt, 1)
            out[t] = np.stack([ctx.seismograms(s, c, p.nt) for s in range(2) for c in range(2)])
    e = rel(out[11], out[0])
    _print("fused vs two-pass f16", e)
    assert np.isfinite(out[11]).all() and e < 1e-2


def test_batch_equals_single_shots():
    _print("case: batch vs solo")
    srcs = [(20 + 13 * k, 6 + k % 3) for k in range(9)]
    p = Problem(_lib.WAVE_PSV, 160, 80, 10, 10.0, 7e-4, 60, seed=11, sources=srcs)
    with p.context() as ctx:
        ctx.tile = 4  # 4 shots per workgroup, 9 shots: ragged last group
        ctx.run(0, p.nt, 1)
        batch = [ctx.get_state(s) for s in range(len(srcs))]
        bseis = [ctx.seismograms(s, 1, p.nt) for s in range(len(srcs))]
    for s, src in enumerate(srcs):
        q = Problem(_lib.WAVE_PSV, 160, 80, 10, 10.0, 7e-4, 60, seed=11, sources=[src])
        with q.context() as ctx:
            ctx.run(0, q.nt, 1)
            assert ctx.get_state(0) == batch[s]
            assert np.array_equal(ctx.seismograms(0, 1, q.nt), bseis[s])


@pytest.mark.parametrize("storage", [_lib.STORE_F32, _lib.STORE_F16])
def test_autotune_preserves_state(storage):
    """Tuning (pass timings and graph-replayed chunks with one, two and
    four chains) leaves the state, the scales and the seismograms as they
    were: the run continued after it equals the untuned run bitwise."""
    _print("case: autotune", storage)
    p = Problem(_lib.WAVE_PSV, 150, 90, 12, 10.0, 7e-4, 50, seed=2, sources=[(70, 6), (40, 9), (100, 4), (75, 20)])
    with p.context(storage=storage) as a, p.context(storage=storage) as b:
        a.run(0, 25, 1)
        b.run(0, 25, 1)
        tile, ms = a.autotune(2)
        _print("best tile", tile, ms, "chains", a.chains)
        a.tile, b.tile = tile, tile  # the same shape, so FP16 rounding is the same
        b.chains = a.chains
        a.run(25, 25, 1)
        b.run(25, 25, 1)
        for s in range(4):
            assert a.get_state(s) == b.get_state(s)
            for c in range(2):
                assert np.array_equal(a.seismograms(s, c, p.nt), b.seismograms(s, c, p.nt))
