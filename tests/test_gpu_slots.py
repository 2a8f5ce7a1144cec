"""Shot slots: a shot that starts while its context runs (fd23_reset_shot,
then fd23_set_shot_slot with its first step) gives the seismograms of the
same shot run alone from step 0, bitwise, on every stepping path: the
two-pass P-SV and SH kernels, both fused P-SV steps and the SH block step.
The slot it takes first ran another shot, so the reset is exercised; an
idle slot beside it is skipped and leaves the other shots unchanged."""
import numpy as np
import pytest

from fd23_amd import _lib
from helpers import Problem

pytestmark = pytest.mark.gpu

ITREC = 3
JOIN = 37          # the step the new shot starts at: odd, not a multiple of ITREC
NT0, NT1 = 100, 90  # steps of shot 0 (from step 0) and of the joining shot


def _frames(n):
    return (n + ITREC - 1) // ITREC


def _tile(which):
    """Tile index of a stepping path, looked up in the library's table when
    the test runs (the library is not loaded while tests are collected: it
    is loaded after torch, as in every GPU test)."""
    import ctypes

    lib = _lib.load()

    def shape(t):
        v = [ctypes.c_int() for _ in range(4)]
        lib.fd23_tile_shape(t, *(ctypes.byref(x) for x in v))
        return [x.value for x in v]  # rz, spb, segment rows or steps per launch, block rows

    if which == "two-pass":
        return 0
    if which == "two shots per workgroup":
        return next(t for t in range(lib.fd23_num_tiles())
                    if lib.fd23_tile_kind(t) == _lib.TILE_TILED and shape(t)[1] > 1)
    kind, pred = {"fused strips": (_lib.TILE_FUSED, lambda t: shape(t)[3] > 0),
                  "fused tiles": (_lib.TILE_FUSED, lambda t: shape(t)[3] == 0),
                  "block": (_lib.TILE_BLOCK, lambda t: shape(t)[2] == 2)}[which]
    return next(t for t in range(lib.fd23_num_tiles()) if lib.fd23_tile_kind(t) == kind and pred(t))


CASES = [
    (_lib.WAVE_PSV, _lib.SRC_P, "two-pass"),
    (_lib.WAVE_PSV, _lib.SRC_LINE, "fused strips"),
    (_lib.WAVE_PSV, _lib.SRC_LINE, "fused tiles"),
    (_lib.WAVE_PSV, _lib.SRC_LINE, "two shots per workgroup"),
    (_lib.WAVE_SH, _lib.SRC_FORCE, "two-pass"),
    (_lib.WAVE_SH, _lib.SRC_FORCE, "block"),
]


def _alone(wave, kind, src, nt, tile):
    p = Problem(wave, 130, 97, 10, 10.0, 7e-4, 130, seed=41, kind=kind, sources=[src])
    ncomp = 2 if wave == _lib.WAVE_PSV else 1
    with p.context(max_frames=60) as ctx:
        ctx.tile = tile
        ctx.run(0, nt, ITREC)
        ctx.sync()
        return [ctx.seismograms(0, c, _frames(nt)) for c in range(ncomp)]


@pytest.mark.parametrize("wave,kind,path", CASES)
def test_shot_joining_a_running_context_equals_its_alone_run(wave, kind, path):
    tile = _tile(path)
    print("case: shot slots", "psv" if wave == _lib.WAVE_PSV else "sh", path, "tile", tile, flush=True)
    a, b, c = (40, 9), (90, 12), (66, 7)  # shot 0, the slot's first occupant, the joining shot
    p = Problem(wave, 130, 97, 10, 10.0, 7e-4, 130, seed=41, kind=kind, sources=[a, b, (20, 20)])
    ncomp = 2 if wave == _lib.WAVE_PSV else 1
    with p.context(max_frames=60) as ctx:
        ctx.tile = tile
        ctx.set_shot_slot(2, active=False)      # an idle slot throughout
        ctx.run(0, JOIN, ITREC)                 # shot 0 and the first occupant of slot 1
        ctx.reset_shot(1)
        ctx.set_shot_slot(1, True, JOIN)
        ctx.set_source(1, kind, c[0], c[1], p.stf, p.M)
        ctx.run(JOIN, 41, ITREC)                # two runs, so a chunk boundary falls inside
        ctx.run(JOIN + 41, NT1 - 41, ITREC)
        ctx.sync()
        got0 = [ctx.seismograms(0, k, _frames(NT0)) for k in range(ncomp)]
        got1 = [ctx.seismograms(1, k, _frames(NT1)) for k in range(ncomp)]
    want0 = _alone(wave, kind, a, NT0, tile)
    want1 = _alone(wave, kind, c, NT1, tile)
    assert any(np.abs(x).max() > 0 for x in want1)
    for k in range(ncomp):
        np.testing.assert_array_equal(got0[k], want0[k])
        np.testing.assert_array_equal(got1[k], want1[k])


def test_slot_arguments_are_checked():
    p = Problem(_lib.WAVE_PSV, 96, 64, 8, 10.0, 7e-4, 20, sources=[(40, 8)])
    lib = _lib.load()
    with p.context() as ctx:
        assert lib.fd23_set_shot_slot(ctx.h, 0, 1, -1) == _lib.EINVAL
        assert lib.fd23_set_shot_slot(ctx.h, 1, 1, 0) == _lib.EINVAL   # one shot only
        assert lib.fd23_reset_shot(ctx.h, 3) == _lib.EINVAL


@pytest.mark.parametrize("storage", [_lib.STORE_F16, _lib.STORE_F32])
def test_a_quieted_slot_stays_zero(storage):
    """The serving engine quiets the slot of a finished request: zero state
    (fd23_reset_shot), no source samples, idle.  On the paths that still
    step idle slots (velocity kernels with two shots per workgroup) its
    state must stay exactly zero, so FP16/BF16 scales, shared by the shots
    of a context, never see it."""
    tile = _tile("two shots per workgroup")
    print("case: quiet slot, storage", storage, "tile", tile, flush=True)
    p = Problem(_lib.WAVE_PSV, 130, 97, 10, 10.0, 7e-4, 130, seed=43, sources=[(40, 9), (90, 12)])
    with p.context(storage=storage, max_frames=60) as ctx:
        ctx.tile = tile
        ctx.run(0, 50, ITREC)              # both shots carry a wavefield
        ctx.reset_shot(1)
        ctx.set_source(1, _lib.SRC_LINE, 0, 0, np.zeros(0), np.zeros(3))
        ctx.set_shot_slot(1, False)
        ctx.run(50, 60, ITREC)
        ctx.sync()
        st1 = np.frombuffer(ctx.get_state(1), np.float32)  # the compute type of both storages
        st0 = np.frombuffer(ctx.get_state(0), np.float32)
    assert np.abs(st0).max() > 0
    assert np.all(st1 == 0)
