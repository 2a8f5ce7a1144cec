"""The copy and tiled-stream calibration kernels compute what their
declarations state: a byte-exact copy, and writes[k] = the reads summed in
order + k over every cell of the grid (FP32 adds, bitwise against numpy
FP32 arithmetic in the same order)."""
import ctypes

import numpy as np
import pytest

from fd23_amd import _lib

pytestmark = pytest.mark.gpu


def test_device_copy_is_exact():
    import torch

    lib = _lib.load()
    g = torch.Generator().manual_seed(3)
    x = torch.randint(-2**31, 2**31 - 1, (4096 + 4,), dtype=torch.int32, generator=g).cuda()
    y = torch.zeros_like(x)
    for grid in (1, 7, 1024):
        y.zero_()
        _lib.check(lib.fd23_device_copy(ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                        x.numel() * 4, grid, None))
        torch.cuda.synchronize()
        assert torch.equal(x, y), grid


@pytest.mark.parametrize("nr,nw,nx,nz", [(1, 1, 64, 16), (8, 3, 192, 48), (7, 2, 128, 32), (3, 1, 320, 16)])
def test_stream_tiles_sum_of_reads(nr, nw, nx, nz):
    import torch

    lib = _lib.load()
    rng = np.random.default_rng(nr * 10 + nw)
    host = rng.standard_normal((nr, nz * nx)).astype(np.float32)
    reads = torch.from_numpy(host).cuda()
    writes = torch.full((nw, nz * nx), np.nan, dtype=torch.float32, device="cuda")
    rd = (ctypes.c_void_p * nr)(*[reads[k].data_ptr() for k in range(nr)])
    wr = (ctypes.c_void_p * nw)(*[writes[k].data_ptr() for k in range(nw)])
    _lib.check(lib.fd23_device_stream_tiles(rd, nr, wr, nw, nx, nz, None))
    torch.cuda.synchronize()
    s = host[0].copy()
    for k in range(1, nr):
        s = (s + host[k]).astype(np.float32)
    got = writes.cpu().numpy()
    for k in range(nw):
        want = (s + np.float32(k)).astype(np.float32)
        assert np.array_equal(got[k].view(np.uint32), want.view(np.uint32)), k
