"""The VALU and MFMA (v_mfma_f32_16x16x4_f32) forms of the hop's link x
half-spinor block product, Y (N x 2k) = U (N x N) H (N x 2k) complex, agree
with an FP64 host product over 1024 blocks (FP32 products: relative error
< 1e-6), for every colour count the library instantiates and RHS counts whose
2k columns fill one to four 16-column MFMA tiles. The same call reports the
rates bench.py compares."""
import ctypes

import pytest
import torch

from mrhs_qcu_amd import _lib
from test_dslash_gpu import stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,k,resident", [(2, 8, 0), (3, 16, 1), (4, 24, 0), (5, 6, 1), (6, 4, 0), (7, 3, 1),
                                          (8, 32, 0), (4, 1, 1)])
def test_link_block_forms_agree(lib, dev, N, k, resident):
    tf = (ctypes.c_double * 2)()
    err = (ctypes.c_double * 2)()
    _lib.check(lib.qcuMrhsLinkBlockBench(N, k, 4, resident, tf, err, stream()))
    print(f"N={N} k={k} resident={resident}: VALU {tf[0]:.2f} / MFMA {tf[1]:.2f} TFLOP/s, "
          f"errors {err[0]:.1e} / {err[1]:.1e}")
    assert err[0] < 1e-6 and err[1] < 1e-6, (err[0], err[1])
    assert tf[0] > 0 and tf[1] > 0


def test_link_block_rejects_bad_sizes(lib, dev):
    tf = (ctypes.c_double * 2)()
    err = (ctypes.c_double * 2)()
    assert lib.qcuMrhsLinkBlockBench(9, 4, 1, 0, tf, err, stream()) != 0
    assert lib.qcuMrhsLinkBlockBench(3, 33, 1, 0, tf, err, stream()) != 0
