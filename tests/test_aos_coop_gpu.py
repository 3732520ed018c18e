"""The user AoS layout read and written by site groups: the qcu.h and
qcuMrhsDslashUser path for 8 or more FP32 / FP16 right-hand sides (a power of
two per launch chunk; 16 for FP16 at N >= 4). Each case is checked against
the FP64 oracle (FP32 <= 1e-6, FP16 <= 3e-3 relative per RHS) and bit for bit
against the kernel in which each lane reads its own records (cache hint bit
4 of the user-layout tile): the two differ only in how records travel, not in
the arithmetic. Extents with an odd x/2 make site groups straddle x rows and
wavefronts start mid-row; 40 RHS run as chunks of 32 and 8."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err, dev_tensor, host_complex, empty_prec

pytestmark = pytest.mark.gpu

TOL = {_lib.SINGLE: 1e-6, _lib.HALF: 3e-3}
DT = {_lib.SINGLE: torch.float32, _lib.HALF: torch.float16}


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def gauge(lib, dev, g, geom, prec):
    gd = dev_tensor(g, dev)
    gn = empty_prec((lib.qcuMrhsGaugeBytes(geom, prec) // _lib.BYTES[prec],), prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), prec, gd.data_ptr(), _lib.DOUBLE, geom, stream()))
    return gn


def spinors(psi, prec, dev, misalign=False):
    """One device buffer per RHS (user layout); misalign: each starts 8 bytes
    past a 16-byte boundary."""
    out = []
    for r in range(psi.shape[0]):
        t = dev_tensor(psi[r].astype(np.complex64), dev).to(DT[prec])
        if misalign:
            pad = 8 // t.element_size()
            big = torch.zeros(t.numel() + pad, dtype=t.dtype, device=dev)
            big[pad:] = t.reshape(-1)
            t = big[pad:].view(t.shape)
        out.append(t)
    return out


def run_user(lib, gn, ins, prec, k, parity, dagger, geom, per_lane):
    t = _lib.tile(1, 256, 0, 1, 0, 16 if per_lane else 0, 0, 0)
    _lib.check(lib.qcuMrhsSetTile(geom, prec | _lib.TILE_USER_LAYOUT, k, ctypes.byref(t)))
    outs = [torch.full_like(x, float("nan")) for x in ins]
    ai, k1 = _lib.ptr_array([x.data_ptr() for x in ins])
    ao, k2 = _lib.ptr_array([x.data_ptr() for x in outs])
    _lib.check(lib.qcuMrhsDslashUser(ao, ai, gn.data_ptr(), prec, k, parity, dagger, geom, stream()))
    torch.cuda.synchronize()
    return outs


def restore(lib, geom, prec, k):
    _lib.check(lib.qcuMrhsSetTile(geom, prec | _lib.TILE_USER_LAYOUT, k, ctypes.byref(_lib.tile(1, 256, 0, 1, 0, 0, 0, 0))))


CASES = [  # (N, prec, k, L, grid)
    (3, _lib.SINGLE, 8, (6, 4, 6, 8), (1, 1, 1, 1)),
    (3, _lib.SINGLE, 16, (6, 4, 6, 8), (1, 1, 1, 2)),
    (3, _lib.SINGLE, 32, (6, 4, 6, 8), (1, 1, 1, 1)),
    (3, _lib.SINGLE, 40, (4, 4, 2, 6), (1, 1, 1, 1)),
    (2, _lib.SINGLE, 16, (6, 4, 6, 8), (1, 2, 1, 1)),
    (4, _lib.SINGLE, 16, (6, 4, 6, 4), (1, 1, 1, 1)),
    (3, _lib.HALF, 16, (6, 4, 6, 8), (1, 1, 1, 1)),
    (2, _lib.HALF, 8, (6, 4, 6, 8), (1, 1, 1, 1)),
    (4, _lib.HALF, 16, (6, 4, 6, 4), (1, 1, 2, 1)),
    (4, _lib.HALF, 8, (6, 4, 6, 4), (1, 1, 1, 1)),  # per-lane by the selection rule
]


@pytest.mark.parametrize("N,prec,k,L,grid", CASES)
def test_group_reads_match_oracle_and_per_lane(lib, dev, N, prec, k, L, grid):
    g = fields.gauge_field(L, N, seed=150 + N, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, N, k, seed=160 + k)
    geom = _lib.geom(L, N, grid)
    gn = gauge(lib, dev, g, geom, prec)
    ins = spinors(psi, prec, dev)
    # the FP16 inputs as stored, for the oracle
    psi_in = np.stack([host_complex(x) for x in ins])
    try:
        for parity, dagger in ((0, 0), (1, 1)):
            got = run_user(lib, gn, ins, prec, k, parity, dagger, geom, per_lane=False)
            ref = run_user(lib, gn, ins, prec, k, parity, dagger, geom, per_lane=True)
            for a, b in zip(got, ref):
                assert torch.equal(a, b), (N, prec, k, parity, dagger)
            err = rel_err(np.stack([host_complex(x) for x in got]),
                          oracle.dslash(g, psi_in, L, parity, dagger, grid=grid))
            assert err < TOL[prec], (N, prec, k, parity, dagger, err)
    finally:
        restore(lib, geom, prec, k)


def test_misaligned_buffers_take_the_per_lane_path(lib, dev):
    """Buffers that are not 16-byte aligned are read element by element."""
    L, N, k, prec = (6, 4, 6, 8), 3, 16, _lib.SINGLE
    g = fields.gauge_field(L, N, seed=171)
    psi = fields.gaussian_spinors(L, N, k, seed=172)
    geom = _lib.geom(L, N)
    gn = gauge(lib, dev, g, geom, prec)
    ins = spinors(psi, prec, dev, misalign=True)
    assert all(x.data_ptr() % 16 == 8 for x in ins)
    got = run_user(lib, gn, ins, prec, k, 1, 0, geom, per_lane=False)
    restore(lib, geom, prec, k)
    err = rel_err(np.stack([host_complex(x) for x in got]), oracle.dslash(g, psi.astype(np.complex64), L, 1, 0))
    assert err < TOL[prec], err


def test_qcu_h_sixteen_fp32_rhs(lib, dev):
    """qcu.h pushBackFermions x 16 + startDslash in FP32 (the group-read
    kernel), inputs overwritten after the pushes, two rounds."""
    L, N, k = (8, 4, 6, 8), 3, 16
    g = fields.gauge_field(L, N, seed=181)
    grid = _lib.QcuGrid()
    param = _lib.QcuParam()
    for d in range(4):
        grid.grid_size[d] = 1
        param.lattice_size[d] = L[d]
    lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
    gd = dev_tensor(g, dev)
    lib.loadQcuGauge(gd.data_ptr())
    lib.getDslash(0, -3.5, N, k, 1)
    for rnd, parity in ((0, 1), (1, 0)):
        psi = fields.gaussian_spinors(L, N, k, seed=182 + rnd)
        ins = [dev_tensor(psi[r].astype(np.complex64), dev) for r in range(k)]
        outs = [torch.full_like(t, float("nan")) for t in ins]
        torch.cuda.synchronize()
        for r in range(k):
            lib.pushBackFermions(outs[r].data_ptr(), ins[r].data_ptr())
        for t in ins:
            t.zero_()
        torch.cuda.synchronize()
        lib.startDslash(parity, rnd)
        got = np.stack([host_complex(t) for t in outs])
        assert rel_err(got, oracle.dslash(g, psi.astype(np.complex64), L, parity, rnd)) < 1e-6
