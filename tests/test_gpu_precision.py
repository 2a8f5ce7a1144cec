"""Reduced-precision storage (FP16 / BF16 with per-field power-of-two
scales, FP32 arithmetic) against the FP64 oracle.

Error measure: relative L2 of all seismograms of a shot against the FP64
oracle.  Stated tolerances: FP32 storage 1e-5, FP16 storage 1e-3, BF16
storage 1e-2 (8 significant bits: one rounding is 2^-9 relative).
"""
import numpy as np
import pytest

from fd23_amd import _lib
from helpers import Problem, rel

pytestmark = pytest.mark.gpu

TOL = {_lib.STORE_F32: 1e-5, _lib.STORE_F16: 1e-3, _lib.STORE_BF16: 1e-2}
NAMES = {_lib.STORE_F32: "f32", _lib.STORE_F16: "f16", _lib.STORE_BF16: "bf16"}

_CACHE = {}


def _reference(wave, kind):
    key = (wave, kind)
    if key not in _CACHE:
        p = Problem(wave, 300, 220, 20, 10.0, 8e-4, 900, seed=4242, kind=kind, f64=True,
                    sources=[(150, 12)], rec_z=40)
        _, seis = p.oracle()
        _CACHE[key] = (p, seis)
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
    print("storage", NAMES[storage], "error vs fp64", errs, flush=True)
    assert max(errs) < TOL[storage]


def test_scales_stay_powers_of_two():
    p, _ = _reference(_lib.WAVE_PSV, _lib.SRC_LINE)
    with p.context(storage=_lib.STORE_F16) as ctx:
        ctx.run(0, 200)
        s, amax = ctx.get_scales(16)
    m, e = np.frexp(s[s > 0])
    assert np.all(m == 0.5), s
    assert np.all(np.isfinite(amax))
