"""FP16 and BF16 storage at the largest configuration's size: P-SV
4096x2048, 8 shots (explosive Ricker 20 Hz at z=8, 2048 receiver pairs at
z=4), over the first 1000 of its 6000 steps.

Reference: the FP64-storage run of the same inputs on the GPU, itself
checked here against the FP64 oracle on shot 0.  Bar: the error of the same
storage emulated on the FP64 oracle (every array a pass updates rounded
after the pass, helpers.emulated_storage), computed for shot 0 at full
size, times 1.3 for shot 0 and times 2 for the other shots (their floors
differ with the source position).  The stated target of 1e-3 relative L2
is printed beside each error: per-array scaled FP16/BF16 storage does not
reach it (its rounding noise accumulates with the step count), and the
test asserts the storage floor, not the target.
"""
import functools

import numpy as np
import pytest

import oracle as O
from fd23_amd import _lib, workloads as W
from helpers import emulated_storage, unpack_cpml

pytestmark = pytest.mark.gpu

NT = 1000
TARGET = 1e-3


def _print(*a):
    print(*a, flush=True)


@functools.lru_cache(maxsize=None)
def _workload(storage):
    # FP64 storage takes the double material, the others the float one
    return W.config_psv_precision(_lib.STORE_F64 if storage == _lib.STORE_F64 else _lib.STORE_F32).build()


class _Shot0:
    """Shot 0 of the workload in the form helpers.emulated_storage reads."""

    def __init__(self, wl):
        self.wave, self.nx, self.nz, self.npml = wl.wave, wl.nx, wl.nz, wl.npml
        self.topabsorb, self.kind, self.nt = wl.topabsorb, wl.kind, NT
        self.mat, self.packed, self.M = wl.mat, wl.packed, wl.M
        self.stf = wl.stf[:NT]
        self.sources = wl.sources[:1]
        self.rx, self.rz = wl.receivers()


@functools.lru_cache(maxsize=None)
def _oracle_f64():
    wl = _workload(_lib.STORE_F64)
    p = _Shot0(wl)
    cp = unpack_cpml(p.packed, p.wave, p.nx, p.nz, np.float64)
    sim = O.PSV(p.nx, p.nz, p.npml, dict(zip(("BU", "BW", "MU", "MUA", "LAM"), p.mat)), cp, p.topabsorb,
                np.float64)
    xs, zs = p.sources[0]
    u, w = O.run_psv(sim, NT, p.stf.astype(np.float64), p.M.astype(np.float64), xs, zs, rx=p.rx, rz=p.rz)
    return np.stack([u, w])


@functools.lru_cache(maxsize=None)
def _floor(fmt):
    emu = np.stack(emulated_storage(_Shot0(_workload(_lib.STORE_F64)), fmt))
    ref = _oracle_f64()
    return float(np.linalg.norm(emu - ref) / np.linalg.norm(ref))


@functools.lru_cache(maxsize=None)
def _gpu(storage):
    wl = _workload(storage)
    wl.storage = storage
    ctx = wl.context(max_frames=NT)
    try:
        ctx.run(0, NT, 1)
        ctx.sync()
        return np.stack([np.stack([ctx.seismograms(s, c, NT).astype(np.float64) for c in range(2)])
                         for s in range(wl.nshots)])
    finally:
        ctx.close()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_fp64_gpu_reference_equals_fp64_oracle():
    _print("case: 4096x2048 FP64 GPU vs FP64 oracle, shot 0, %d steps" % NT)
    g = _gpu(_lib.STORE_F64)
    err = _rel(g[0], _oracle_f64())
    _print("fp64 gpu vs oracle", err)
    assert np.isfinite(g).all() and err < 1e-9


@pytest.mark.parametrize("storage,fmt", [(_lib.STORE_F16, "f16"), (_lib.STORE_BF16, "bf16")])
def test_half_storage_at_its_floor(storage, fmt):
    _print("case: 4096x2048 %s vs FP64, 8 shots, %d steps" % (fmt, NT))
    ref = _gpu(_lib.STORE_F64)
    got = _gpu(storage)
    assert np.isfinite(got).all()
    errs = [_rel(got[s], ref[s]) for s in range(got.shape[0])]
    floor = _floor(fmt)
    _print("%s error vs fp64 per shot" % fmt, ["%.3e" % e for e in errs])
    _print("%s emulated floor (shot 0) %.3e; stated target %.0e: %s" % (
        fmt, floor, TARGET, "met" if max(errs) <= TARGET else "not met (storage floor above it)"))
    assert errs[0] <= 1.3 * floor
    assert max(errs) <= 2.0 * floor
