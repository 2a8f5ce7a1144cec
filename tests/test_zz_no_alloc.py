"""fd23_set_model_noalloc (the entry point that allocates no device memory)
against fd23_set_model on the same inputs: every material array bitwise and
the same stability number, on consecutive calls whose second input is
smaller than the first (the scratch keeps the previous call's contents, so
stale staging or reduction keys would show), across P-SV FP32, P-SV FP64
and SH; a model above the scratch capacity is rejected with FD23_EINVAL and
the next call still succeeds."""
import numpy as np
import pytest

from fd23_amd import _lib
from fd23_amd.core import Context
from helpers import layered_model

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _model(nx, nz, seed, vscale=1.0, f64=False):
    vp, vs, den = layered_model(nx, nz, seed)
    vp = vp * np.float32(vscale)
    vs = vs * np.float32(vscale)
    vs[:, :2] = 0.0  # water in the left edge band
    if f64:
        vp, vs, den = (a.astype(np.float64) for a in (vp, vs, den))
    return vp, vs, den


def _both(wave, storage, nx, nz, model, h, dt):
    nmat = 5 if wave == _lib.WAVE_PSV else 3
    with Context(wave, nx, nz, 8, storage=storage) as a:
        sa = a.set_model(*model, h, dt)
        ma = [a.material(i) for i in range(nmat)]
    with Context(wave, nx, nz, 8, storage=storage) as b:
        sb = b.set_model(*model, h, dt, noalloc=True)
        mb = [b.material(i) for i in range(nmat)]
    return sa, ma, sb, mb


# each case's model is smaller and slower than the previous one's, so
# leftover staging or keys from the call before would change the result
CASES = [
    (_lib.WAVE_PSV, _lib.STORE_F32, 421, 263, 1.3),
    (_lib.WAVE_PSV, _lib.STORE_F32, 150, 90, 1.0),
    (_lib.WAVE_PSV, _lib.STORE_F64, 131, 77, 0.9),
    (_lib.WAVE_SH, _lib.STORE_F32, 97, 41, 0.8),
]


def test_noalloc_equals_set_model_on_consecutive_calls():
    h, dt = 10.0, 5e-4
    for wave, storage, nx, nz, vscale in CASES:
        _print("case: set_model_noalloc", wave, storage, nx, nz, vscale)
        model = _model(nx, nz, nx + nz, vscale, f64=storage == _lib.STORE_F64)
        sa, ma, sb, mb = _both(wave, storage, nx, nz, model, h, dt)
        _print("stab", sa, sb)
        assert sa == sb
        for i, (x, y) in enumerate(zip(ma, mb)):
            assert x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8)), (wave, storage, i)


def test_noalloc_rejects_a_model_above_the_scratch():
    lib = _lib.load()
    cap = lib.fd23_set_model_noalloc_capacity()
    assert cap == 32 << 20
    nx, nz = 2048, 1400  # 3 FP32 arrays of 2048 x 1400 cells: 34.4 MB > 32 MiB
    assert 3 * 4 * nx * nz > cap
    _print("case: set_model_noalloc over capacity", nx, nz)
    vp = np.full((nz, nx), 3000.0, np.float32)
    vs = vp / np.float32(1.732)
    den = np.full((nz, nx), 2400.0, np.float32)
    with Context(_lib.WAVE_PSV, nx, nz, 8) as ctx:
        with pytest.raises(_lib.Fd23Error) as e:
            ctx.set_model(vp, vs, den, 10.0, 5e-4, noalloc=True)
        assert "scratch" in str(e.value)
    # the next call on the same GPU works
    wave, storage, nx, nz, vscale = CASES[1]
    model = _model(nx, nz, 7, vscale)
    sa, ma, sb, mb = _both(wave, storage, nx, nz, model, 10.0, 5e-4)
    assert sa == sb and all(np.array_equal(x, y) for x, y in zip(ma, mb))
