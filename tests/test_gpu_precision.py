"""Reduced-precision storage (FP16 / BF16 with per-field power-of-two
scales, FP32 arithmetic) against the FP64 oracle.

Error measure: relative L2 of all seismograms of a shot against the FP64
oracle.  FP32 storage: 1e-5.  FP16 and BF16: the rounding of the stored
arrays is itself the error (one rounding is 2^-11 relative for FP16, 2^-8
for BF16, and every array is rounded once per step), so the bar is the
error of the same storage emulated on the FP64 oracle (helpers.
emulated_storage), times 1.3.  On this 300x220, 900-step case that floor is
4.8e-3 (FP16) and 3.8e-2 (BF16) for P-SV.
"""
import numpy as np
import pytest

from fd23_amd import _lib
from helpers import Problem, emulated_storage, rel

pytestmark = pytest.mark.gpu

NAMES = {_lib.STORE_F32: "f32", _lib.STORE_F16: "f16", _lib.STORE_BF16: "bf16"}
FLOOR_FACTOR = 1.3

_CACHE = {}


def _reference(wave, kind):
    key = (wave, kind)
    if key not in _CACHE:
        p = Problem(wave, 300, 220, 20, 10.0, 8e-4, 900, seed=4242, kind=kind, f64=True,
                    sources=[(150, 12)], rec_z=40)
        _, seis = p.oracle()
        _CACHE[key] = (p, seis)
    return _CACHE[key]


def _floor(wave, kind, fmt):
    key = (wave, kind, fmt)
    if key not in _CACHE:
        p, ref = _reference(wave, kind)
        emu = emulated_storage(p, fmt)
        _CACHE[key] = max(rel(e, r) for e, r in zip(emu, ref))
    return _CACHE[key]


@pytest.mark.parametrize("storage", [_lib.STORE_F32, _lib.STORE_F16, _lib.STORE_BF16])
@pytest.mark.parametrize("wave,kind", [(_lib.WAVE_PSV, _lib.SRC_LINE), (_lib.WAVE_SH, _lib.SRC_FORCE)])
def test_storage_vs_fp64(storage, wave, kind):
    print("case: storage=%s wave=%d kind=%d" % (NAMES[storage], wave, kind), flush=True)
    p, ref = _reference(wave, kind)
    with p.context(storage=storage) as ctx:
        ctx.run(0, p.nt)
        ctx.sync()
        errs = [rel(ctx.seismograms(0, c, p.nt), ref[c]) for c in range(len(ref))]
        fields = ("U", "W") if wave == _lib.WAVE_PSV else ("V",)
        assert all(np.isfinite(ctx.field(f)).all() for f in fields)
    if storage == _lib.STORE_F32:
        print("storage f32 error vs fp64", errs, flush=True)
        assert max(errs) < 1e-5
        return
    floor = _floor(wave, kind, NAMES[storage])
    print("storage", NAMES[storage], "error vs fp64", errs, "emulated floor", floor, flush=True)
    assert max(errs) <= FLOOR_FACTOR * floor


def test_scales_stay_powers_of_two():
    p, _ = _reference(_lib.WAVE_PSV, _lib.SRC_LINE)
    with p.context(storage=_lib.STORE_F16) as ctx:
        ctx.run(0, 200)
        s, amax = ctx.get_scales(16)
    m, e = np.frexp(s[s > 0])
    assert np.all(m == 0.5), s
    assert np.all(np.isfinite(amax))
