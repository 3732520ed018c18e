"""The reference's OWN kernels (built from its sources into oracle/_ref/)
against the CPU oracle and the library, on identical inputs (GPU)."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err

pytestmark = pytest.mark.gpu


def ref_lib():
    r = oracle.ref_kernels()
    if r is None:
        pytest.skip("oracle/_ref/libqcu_ref.so not built (reference sources absent when building)")
    return r


def run_ref(ref, dev, g, psi, L, N, parity, dagger, grid, prec=2):
    k = psi.shape[0]
    S = fields.half_volume(L)
    ft = np.float64 if prec == 2 else np.float32
    gd = torch.from_numpy(oracle.shift_gauge(g.reshape(-1), True, S, N).view(ft).copy()).to(dev)
    ins = [torch.from_numpy(oracle.shift_vector(psi[r].reshape(-1), True, S, N).view(ft).copy()).to(dev)
           for r in range(k)]
    outs = [torch.empty_like(t) for t in ins]
    arr = ctypes.c_void_p * k
    ai, ao = arr(*[t.data_ptr() for t in ins]), arr(*[t.data_ptr() for t in outs])
    rc = ref.ref_dslash(ctypes.cast(ao, ctypes.POINTER(ctypes.c_void_p)),
                        ctypes.cast(ai, ctypes.POINTER(ctypes.c_void_p)), gd.data_ptr(), k, N, prec,
                        (ctypes.c_int * 4)(*L), (ctypes.c_int * 4)(*grid), parity, dagger)
    assert rc == 0
    ct = np.complex128 if prec == 2 else np.complex64
    return np.stack([oracle.shift_vector(t.cpu().numpy().view(ct), False, S, N).reshape(S, 4, N) for t in outs])


@pytest.mark.parametrize("N", [3, 4, 5])
@pytest.mark.parametrize("k", [1, 2, 5])
def test_reference_kernel_equals_oracle(dev, N, k):
    ref = ref_lib()
    L = (4, 4, 4, 8)
    g = fields.gauge_field(L, N, seed=200 + N)
    psi = fields.gaussian_spinors(L, N, k, seed=300 + k)
    for parity in (0, 1):
        for dagger in (0, 1):
            for grid in ((1, 1, 1, 1), (2, 1, 2, 2)):
                got = run_ref(ref, dev, g, psi, L, N, parity, dagger, grid)
                want = oracle.dslash(g, psi, L, parity, dagger, grid=grid)
                assert rel_err(got, want) < 1e-13, (parity, dagger, grid)


def test_reference_kernel_config1_vs_library(dev, lib):
    """8^4 SU(3) nrhs = 1 FP64: the reference's dslash and the library's agree."""
    ref = ref_lib()
    L, N = (8, 8, 8, 8), 3
    g = fields.gauge_field(L, N, seed=0)
    psi = fields.gaussian_spinors(L, N, 1, seed=0)
    want = run_ref(ref, dev, g, psi, L, N, 0, 0, (1, 1, 1, 1))
    from test_dslash_gpu import run_packed

    got = run_packed(lib, dev, g, psi, L, N, 0, 0, _lib.DOUBLE)
    assert rel_err(got, want) < 1e-13


def test_reference_shift_kernels_equal_index_maps(dev):
    ref = ref_lib()
    L, N = (4, 4, 4, 4), 3
    S = fields.half_volume(L)
    v = fields.gaussian_spinors(L, N, 1, seed=7)[0].reshape(-1)
    g = fields.gauge_field(L, N, seed=8).reshape(-1)
    vd = torch.from_numpy(v.view(np.float64).copy()).to(dev)
    gd = torch.from_numpy(g.view(np.float64).copy()).to(dev)
    vs, gs = torch.empty_like(vd), torch.empty_like(gd)
    assert ref.ref_shift_vector(vs.data_ptr(), vd.data_ptr(), 0, L[0] // 2, L[1], L[2], L[3]) == 0
    assert ref.ref_shift_gauge(gs.data_ptr(), gd.data_ptr(), 0, L[0], L[1], L[2], L[3]) == 0
    assert np.array_equal(vs.cpu().numpy().view(np.complex128), oracle.shift_vector(v, True, S, N))
    assert np.array_equal(gs.cpu().numpy().view(np.complex128), oracle.shift_gauge(g, True, S, N))
