"""Eight-real SU(3) links (qcuMrhsPackGauge8 + qcuMrhsDslash8) against the
FP64 oracle on the full links: FP64 <= 1e-13, FP32 <= 1e-6, FP16 / BF16
storage stated and bounded (3e-3 / 2e-2). Every link is rebuilt from U01,
U02, U10 and the phases of U00 and U20 while it is staged in LDS; the
antiperiodic-time sign (det = -1 on the last time slice) is exercised."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err, dev_tensor, host_complex, empty_prec

pytestmark = pytest.mark.gpu

TOL = {2: 1e-13, 1: 1e-6, 0: 3e-3, 4: 2e-2}


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def pack8(lib, dev, g_user, L, prec, t_sign):
    geom = _lib.geom(L, 3)
    gd = dev_tensor(g_user, dev)
    nb = lib.qcuMrhsGauge8Bytes(geom, prec)
    assert nb == fields.half_volume(L) * 2 * 8 * 4 * _lib.BYTES[prec]
    g8 = empty_prec((nb // _lib.BYTES[prec],), prec, dev)
    rc = lib.qcuMrhsPackGauge8(g8.data_ptr(), prec, gd.data_ptr(), _lib.DOUBLE, geom, t_sign, stream())
    return rc, g8


def run8(lib, dev, g_user, psi, L, parity, dagger, prec, t_sign, tile=None, grid=(1, 1, 1, 1)):
    k = psi.shape[0]
    rc, g8 = pack8(lib, dev, g_user, L, prec, t_sign)
    _lib.check(rc)
    geom = _lib.geom(L, 3, grid=grid)
    ins = [dev_tensor(psi[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    pin = empty_prec(psi.shape[1:] + (k,), prec, dev)
    _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), prec, arr, _lib.DOUBLE, k, geom, stream()))
    pout = torch.empty_like(pin)
    tp = ctypes.byref(tile) if tile is not None else None
    _lib.check(lib.qcuMrhsDslash8(pout.data_ptr(), pin.data_ptr(), g8.data_ptr(), prec, k, parity, dagger, t_sign,
                                  geom, tp, stream()))
    outs = [torch.empty(psi.shape[1:] + (2,), dtype=torch.float64, device=dev) for _ in range(k)]
    arr2, keep2 = _lib.ptr_array([t.data_ptr() for t in outs])
    _lib.check(lib.qcuMrhsUnpackSpinors(arr2, _lib.DOUBLE, pout.data_ptr(), prec, k, geom, stream()))
    torch.cuda.synchronize()
    return np.stack([host_complex(t) for t in outs])


def test_fp64_antiperiodic_8x8x8x8(lib, dev):
    L = (8, 8, 8, 8)
    g = fields.gauge_field(L, 3, seed=0)
    psi = fields.gaussian_spinors(L, 3, 4, seed=1)
    for parity in (0, 1):
        for dagger in (0, 1):
            got = run8(lib, dev, g, psi, L, parity, dagger, _lib.DOUBLE, -1, tile=_lib.tile(1, 128, 1))
            err = rel_err(got, oracle.dslash(g, psi, L, parity, dagger))
            print(f"fp64 eight-real parity {parity} dagger {dagger}: {err:.3e}")
            assert err < TOL[2]


def test_fp64_periodic_and_boundary_skip(lib, dev):
    """Periodic links (t_sign = +1) and split dimensions (boundary terms skipped)."""
    L = (8, 4, 4, 8)
    gp = fields.gauge_field(L, 3, seed=3, antiperiodic_t=False)
    psi = fields.gaussian_spinors(L, 3, 4, seed=4)
    got = run8(lib, dev, gp, psi, L, 1, 0, _lib.DOUBLE, 1, tile=_lib.tile(2, 64, 1))
    assert rel_err(got, oracle.dslash(gp, psi, L, 1, 0)) < TOL[2]
    ga = fields.gauge_field(L, 3, seed=3)
    for grid in ((1, 1, 1, 2), (2, 2, 2, 2)):
        got = run8(lib, dev, ga, psi, L, 0, 1, _lib.DOUBLE, -1, tile=_lib.tile(1, 64, 1), grid=grid)
        assert rel_err(got, oracle.dslash(ga, psi, L, 0, 1, grid=grid)) < TOL[2]


@pytest.mark.parametrize("tile", [None, (1, 128, 1), (2, 128, 1), (4, 256, 1), (2, 128, 1, 2, 4, 0, 0, 1),
                                  (2, 64, 1, 2, 4, 8, 0, 1)])
def test_fp32_tiles(lib, dev, tile):
    L, k = (8, 8, 8, 16), 16
    g = fields.gauge_field(L, 3, seed=0, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, 3, k, seed=2)
    t = _lib.tile(*tile) if tile else None
    for parity, dagger in ((0, 0), (1, 1)):
        got = run8(lib, dev, g, psi, L, parity, dagger, _lib.SINGLE, -1, tile=t)
        err = rel_err(got, oracle.dslash(g, psi, L, parity, dagger))
        print(f"fp32 eight-real tile {tile} parity {parity}: {err:.3e}")
        assert err < TOL[1]


@pytest.mark.parametrize("prec", [_lib.HALF, _lib.BF16])
def test_16bit_storage(lib, dev, prec):
    L, k = (8, 8, 8, 16), 16
    g = fields.gauge_field(L, 3, seed=0, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, 3, k, seed=2)
    got = run8(lib, dev, g, psi, L, 0, 0, prec, -1)
    err = rel_err(got, oracle.dslash(g, psi, L, 0, 0))
    print(f"eight-real 16-bit storage prec {prec}: {err:.3e}")
    assert err < TOL[prec]


def test_pack_and_launch_refusals(lib, dev):
    L = (4, 4, 4, 4)
    ga = fields.gauge_field(L, 3, seed=0)  # antiperiodic: det = -1 on the last slice
    rc, _ = pack8(lib, dev, ga, L, _lib.SINGLE, 1)
    assert rc != 0 and "SU(3)" in lib.qcuMrhsLastError().decode()
    rc, g8 = pack8(lib, dev, ga, L, _lib.SINGLE, -1)
    assert rc == 0
    unit = np.zeros_like(ga)
    unit[..., 0, 0] = unit[..., 1, 1] = unit[..., 2, 2] = 1.0
    rc, _ = pack8(lib, dev, unit, L, _lib.SINGLE, 1)
    assert rc != 0 and "1e-6" in lib.qcuMrhsLastError().decode()
    # links staged in LDS are required: a tile without LDS staging is refused
    geom = _lib.geom(L, 3)
    k = 4
    pin = torch.zeros(fields.half_volume(L) * 12 * k * 2, device=dev)
    pout = torch.empty_like(pin)
    rc = lib.qcuMrhsDslash8(pout.data_ptr(), pin.data_ptr(), g8.data_ptr(), _lib.SINGLE, k, 0, 0, -1, geom,
                            ctypes.byref(_lib.tile(1, 128, 0)), stream())
    assert rc != 0 and "link_lds" in lib.qcuMrhsLastError().decode()
