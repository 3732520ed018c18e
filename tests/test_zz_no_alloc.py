"""qcuMrhsDslashExNoAlloc (the allocation-free variant of qcuMrhsDslashEx:
norm partials in a fixed device array instead of a per-call allocation)
against qcuMrhsDslashEx on the same inputs: outputs bit-equal and per-RHS
norms equal, on two calls in a row, the second with a smaller lattice and
fewer right-hand sides (the shared array keeps the first call's contents);
sizes beyond its fixed scratch are rejected with a message."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from helpers import dev_tensor, empty_prec

pytestmark = pytest.mark.gpu

def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def entry(lib):
    return lib.qcuMrhsDslashExNoAlloc


def one_case(lib, dev, L, N, k, prec, use_x, seed):
    geom = _lib.geom(L, N)
    g = fields.gauge_field(L, N, seed=seed)
    psi = fields.gaussian_spinors(L, N, k, seed=seed + 1)
    x = fields.gaussian_spinors(L, N, k, seed=seed + 2)
    conv = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, prec) // _lib.BYTES[prec],), prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), prec, dev_tensor(g, dev).data_ptr(), _lib.DOUBLE, geom,
                                    stream()))
    pin = dev_tensor(fields.to_packed(psi), dev).to(conv[prec])
    px = dev_tensor(fields.to_packed(x), dev).to(conv[prec]) if use_x else None
    sc = torch.from_numpy((2.0 ** np.arange(-2, k - 2) % 7 + 0.5).astype(np.float32)).to(dev)
    outs, norms = [], []
    for fn in (lib.qcuMrhsDslashEx, entry(lib)):
        o = torch.full_like(pin, float("nan"))
        nrm = torch.full((k,), float("nan"), dtype=torch.float64, device=dev)
        _lib.check(fn(o.data_ptr(), prec, pin.data_ptr(), px.data_ptr() if use_x else None, prec, 0.5, -0.25,
                      sc.data_ptr(), sc.data_ptr() if use_x else None, sc.data_ptr(), nrm.data_ptr(),
                      gnat.data_ptr(), k, 1, 0, geom, stream()))
        torch.cuda.synchronize()
        outs.append(o)
        norms.append(nrm)
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(norms[0], norms[1])
    assert torch.isfinite(norms[1]).all()


def test_no_alloc_matches_dslash_ex(lib, dev):
    one_case(lib, dev, (8, 8, 8, 16), 3, 16, _lib.SINGLE, True, 80)
    one_case(lib, dev, (4, 4, 4, 8), 3, 6, _lib.SINGLE, False, 90)
    one_case(lib, dev, (4, 4, 4, 8), 4, 8, _lib.HALF, True, 100)
    one_case(lib, dev, (4, 4, 4, 4), 2, 2, _lib.DOUBLE, False, 110)


def test_no_alloc_rejects_oversize(lib, dev):
    """48^3 x 96 with 32 RHS needs more norm scratch than the fixed array."""
    L, N, k, prec = (48, 48, 48, 96), 3, 32, _lib.SINGLE
    geom = _lib.geom(L, N)
    nrm = torch.zeros(k, dtype=torch.float64, device=dev)
    dummy = torch.zeros(16, dtype=torch.float32, device=dev)
    rc = entry(lib)(dummy.data_ptr(), prec, dummy.data_ptr(), None, prec, 0.0, 1.0, None, None, None,
                    nrm.data_ptr(), dummy.data_ptr(), k, 0, 0, geom, stream())
    assert rc != 0
    assert "scratch" in lib.qcuMrhsLastError().decode()
