"""The streaming kernels behind the bandwidth ceiling and the qcu.h push
measurements: each mode moves exactly the bytes it claims (copy, in-place
rewrite, strided copy into a 16-RHS packed slot), rejects bad arguments, and
reports a positive rate."""
import ctypes

import pytest
import torch

from mrhs_qcu_amd import _lib
from test_dslash_gpu import stream

pytestmark = pytest.mark.gpu


def run(lib, mode, dst, src, nbytes, reps=2):
    gbs = ctypes.c_double()
    _lib.check(lib.qcuMrhsStreamBandwidth(mode, dst.data_ptr() if dst is not None else None,
                                          src.data_ptr() if src is not None else None, nbytes, reps,
                                          ctypes.byref(gbs), stream()))
    torch.cuda.synchronize()
    assert gbs.value > 0
    return gbs.value


def test_stream_modes_move_the_right_bytes(lib, dev):
    n = 1 << 20  # floats (4 MiB)
    src = torch.randn(n, device=dev)
    dst = torch.zeros(n, device=dev)
    run(lib, 2, dst, src, n * 4)
    assert torch.equal(dst, src)
    run(lib, 3, dst, None, n * 4)  # dst = dst * 1 in place
    assert torch.equal(dst, src)
    run(lib, 1, dst, None, n * 4)  # write ones
    assert torch.equal(dst, torch.ones_like(dst))
    run(lib, 0, dst, src, n * 4)  # read only: dst untouched for finite input
    assert torch.equal(dst, torch.ones_like(dst))


def test_stream_strided_copy_into_packed_slot(lib, dev):
    ne = 3 * 4096 + 2  # 8-byte elements (an odd count of 16-byte chunks)
    src = torch.randn(ne * 2, device=dev)
    dst = torch.full((ne * 2 * 16,), float("nan"), device=dev)
    run(lib, 4, dst, src, ne * 8)
    d = dst.view(ne, 16, 2)
    assert torch.equal(d[:, 0, :], src.view(ne, 2))
    assert torch.isnan(d[:, 1:, :]).all()  # the other 15 slots of every line untouched


def test_stream_rejects_bad_arguments(lib, dev):
    a = torch.zeros(64, device=dev)
    gbs = ctypes.c_double()
    assert lib.qcuMrhsStreamBandwidth(5, a.data_ptr(), a.data_ptr(), 256, 1, ctypes.byref(gbs), stream()) != 0
    assert lib.qcuMrhsStreamBandwidth(2, a.data_ptr(), a.data_ptr(), 24, 1, ctypes.byref(gbs), stream()) != 0
    assert lib.qcuMrhsStreamBandwidth(4, a.data_ptr(), None, 256, 1, ctypes.byref(gbs), stream()) != 0
