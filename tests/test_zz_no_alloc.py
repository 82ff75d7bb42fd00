"""pxl_scaled_sum_noalloc (the loss-value sum with its scratch in the library's own device memory,
for callers that cannot allocate) against pxl_scaled_sum (the same kernel with a caller's work
buffer) through the C ABI: the same bits on two calls in a row, the second on a smaller input,
and again after a call of the other entry point."""
import math

import numpy as np
import pytest
import torch

from plenoxels_amd import _lib
from util import say

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _with_work(L, vd, n, scale, work):
    out = torch.full((), -1.0, device=DEV)
    _lib.check(L.pxl_scaled_sum(_lib.ptr(vd), n, scale, _lib.ptr(out), _lib.ptr(work), _lib.stream_handle(DEV)))
    return out


def _noalloc(L, vd, n, scale):
    out = torch.full((), -1.0, device=DEV)
    _lib.check(L.pxl_scaled_sum_noalloc(_lib.ptr(vd), n, scale, _lib.ptr(out), _lib.stream_handle(DEV)))
    return out


def test_noalloc_sum_equals_the_work_buffer_entry_point():
    L = _lib.lib()
    assert L.pxl_scaled_sum_noalloc_scratch_bytes() == 1024
    work = torch.zeros(int(L.pxl_scaled_sum_work_bytes()), dtype=torch.uint8, device=DEV)
    g = torch.Generator().manual_seed(77)
    sizes = [(1 << 20) + 3, 4099, 0, 5]  # a large input, then smaller ones (ragged tails, empty)
    for n in sizes:
        v = (torch.rand(n, generator=g, dtype=torch.float64) * 3).float()
        vd = v.to(DEV)
        scale = 1.0 / (4.0 * max(n, 1))
        say("noalloc sum n", n)
        a = _noalloc(L, vd, n, scale)
        b = _with_work(L, vd, n, scale, work)
        c = _noalloc(L, vd, n, scale)  # after the other entry point: the scratch kept its state right
        torch.cuda.synchronize()
        assert a.item() == b.item() == c.item(), (n, a.item(), b.item(), c.item())
        want = float(np.float32(scale * math.fsum(v.double().tolist())))
        assert abs(a.item() - want) <= 2.0 ** -23 * abs(want), (a.item(), want)


def test_noalloc_rejects_bad_arguments_before_any_launch():
    L = _lib.lib()
    out = torch.empty((), device=DEV)
    assert L.pxl_scaled_sum_noalloc(None, -1, 1.0, _lib.ptr(out), None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc(None, 4, 1.0, _lib.ptr(out), None) == _lib.PXL_E_INVALID
    v = torch.zeros(8, device=DEV)
    assert L.pxl_scaled_sum_noalloc(_lib.ptr(v), 8, 1.0, None, None) == _lib.PXL_E_INVALID
    assert L.pxl_scaled_sum_noalloc(ctypes_offset(v, 4), 4, 1.0, _lib.ptr(out), None) == _lib.PXL_E_INVALID


def ctypes_offset(t, nbytes):
    import ctypes
    return ctypes.c_void_p(t.data_ptr() + nbytes)
