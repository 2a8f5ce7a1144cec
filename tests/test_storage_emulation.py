"""The reduced-precision storage emulation used as the FP16/BF16 bar:
its rounding matches library conversions, FP32 storage stays at FP32
rounding level, and FP16 storage lands at the expected floor."""
import numpy as np
import torch

from fd23_amd import _lib
from helpers import Problem, emulated_storage, rel, round_bf16, round_storage


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(4096) * 10.0 ** rng.integers(-6, 6, 4096),
                        [1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0]])
    x32 = x.astype(np.float32)
    want = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(round_bf16(x32.astype(np.float64)), want)
    # ties go to the even mantissa
    assert round_bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert round_bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


def test_f16_rounding_keeps_scale_and_ulp():
    a = np.linspace(-3e-7, 5e-7, 1001)
    b = a.copy()
    round_storage(b, "f16")
    # largest value scaled into [2^8, 2^9): relative rounding at most 2^-11
    big = np.abs(a) > 1e-9
    assert np.max(np.abs(b[big] - a[big]) / np.abs(a[big])) <= 2.0 ** -11 * (1 + 1e-12)
    assert np.array_equal(np.sign(b[big]), np.sign(a[big]))


def test_emulated_storage_floors():
    p = Problem(_lib.WAVE_PSV, 120, 100, 12, 10.0, 8e-4, 200, seed=11, kind=_lib.SRC_LINE, f64=True,
                sources=[(60, 10)], rec_z=30)
    _, ref = p.oracle()
    err = {fmt: max(rel(e, r) for e, r in zip(emulated_storage(p, fmt), ref)) for fmt in ("f32", "f16", "bf16")}
    print(err)
    assert err["f32"] < 1e-6
    # one rounding per array and step: FP16 well above FP32, BF16 about 8x FP16
    assert 1e-4 < err["f16"] < 2e-2
    assert 4 * err["f16"] < err["bf16"] < 0.2


def test_c_rounding_equals_python_rounding():
    """The oracle's C storage rounding (used for large grids) equals the
    numpy rounding above bitwise, FP16 subnormal range included."""
    import oracle as O

    rng = np.random.default_rng(5)
    for fmt in ("f32", "f16", "bf16"):
        for scale in (1e-9, 3.7, 2.0 ** 20):
            a = rng.standard_normal(20000) * scale * 10.0 ** rng.integers(-12, 1, 20000)
            a[:5] = [0.0, scale, -scale, scale * 2.0 ** -30, scale * (1 + 2.0 ** -11)]
            b = a.copy()
            c = a.copy()
            round_storage(b, fmt)
            O.round_storage(c, fmt)
            assert np.array_equal(b, c), (fmt, scale, np.flatnonzero(b != c)[:5])


def test_block16_rounding_properties():
    """Block-scaled int16 emulation (one power-of-two quantum per 64
    values): every block's values become integer multiples of a quantum
    q with the block maximum in [2^14, 2^15) q, the error is at most q/2,
    values already on that grid are kept, rounding twice changes nothing,
    and blocks are independent (a large value does not coarsen its
    neighbours' blocks)."""
    import oracle as O

    rng = np.random.default_rng(8)
    a = rng.standard_normal(64 * 50) * 10.0 ** np.repeat(rng.integers(-30, 5, 50), 64)
    a[64 * 7:64 * 8] = 0.0                      # an all-zero block
    a[64 * 9:64 * 10] *= 1e-290                 # beneath any exponent range: flushed
    b = a.copy()
    O.round_storage(b, "blk16")
    for k in range(50):
        x, y = a[64 * k:64 * (k + 1)], b[64 * k:64 * (k + 1)]
        m = np.abs(x).max()
        if m == 0.0 or m < 1e-280:
            assert not y.any()
            continue
        q = 2.0 ** (np.floor(np.log2(m)) - 14)
        assert np.all(np.abs(y - x) <= q / 2 * (1 + 1e-12))
        assert np.array_equal(y / q, np.round(y / q)) and np.abs(y / q).max() <= 32767
        assert 2 ** 14 <= np.abs(y).max() / q <= 2 ** 15
    c = b.copy()
    O.round_storage(c, "blk16")
    assert np.array_equal(b, c)
    # a ragged tail block (n not a multiple of 64)
    t = a[:100].copy()
    O.round_storage(t, "blk16")
    assert np.array_equal(t[:64], b[:64]) and np.all(np.abs(t[64:] - a[64:100]) <= np.abs(a[64:100]).max() * 2 ** -14)
    # the grid of integers times 2^-3 with the largest at 20000: kept exactly
    g = np.arange(-32, 32, dtype=np.float64) * 2.0 ** -3
    g[0] = 20000 * 2.0 ** -3
    h = g.copy()
    O.round_storage(h, "blk16")
    assert np.array_equal(g, h)
