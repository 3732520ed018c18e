"""GPU clover term against the textbook restatement (helpers.textbook_clover,
pinned by tests/test_clover_cpu.py): the field (FP64 <= 1e-13), its blockwise
inverse (A A^-1 = 1), gauge covariance of the GPU field, and the application
to packed multi-RHS spinors (FP64 <= 1e-13, FP32 <= 1e-6, FP16 <= 3e-3,
BF16 <= 2e-2 against the FP64 product)."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from helpers import (rel_err, dev_tensor, host_complex, empty_prec, textbook_clover, clover_blocks,
                     gauge_transform)

pytestmark = pytest.mark.gpu

TOL = {2: 1e-13, 1: 1e-6, 0: 3e-3, 4: 2e-2}


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def gpu_clover(lib, dev, g_user, L, N, coeff, prec=_lib.DOUBLE, inverse=False):
    geom = _lib.geom(L, N)
    nb = lib.qcuMrhsCloverBytes(geom, prec)
    S = fields.half_volume(L)
    assert nb == 2 * S * 8 * N * N * _lib.BYTES[prec]
    gd = dev_tensor(g_user, dev)
    a = empty_prec((2, S, 2, 2 * N, 2 * N), prec, dev)
    ai = empty_prec((2, S, 2, 2 * N, 2 * N), prec, dev) if inverse else None
    _lib.check(lib.qcuMrhsCloverCompute(a.data_ptr(), ai.data_ptr() if inverse else None, prec, gd.data_ptr(),
                                        _lib.DOUBLE, coeff, geom, stream()))
    return a, ai


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_field_matches_textbook(lib, dev, N):
    L = (4, 4, 4, 8)
    g = fields.gauge_field(L, N, seed=N)
    a, _ = gpu_clover(lib, dev, g, L, N, -0.7)
    got = host_complex(a)
    want = clover_blocks(textbook_clover(g, L, -0.7), N)
    err = np.abs(got - want).max()
    print(f"clover field N={N}: max abs err {err:.3e}")
    assert err < 1e-13


def test_inverse_and_hermiticity(lib, dev):
    L, N = (4, 4, 4, 8), 3
    g = fields.gauge_field(L, N, seed=7)
    a, ai = gpu_clover(lib, dev, g, L, N, 0.5, inverse=True)
    A, Ai = host_complex(a), host_complex(ai)
    assert np.abs(A - np.conj(np.swapaxes(A, -1, -2))).max() < 1e-15
    eye = np.eye(2 * N)
    assert np.abs(A @ Ai - eye).max() < 1e-12
    assert np.abs(Ai @ A - eye).max() < 1e-12


def test_gpu_field_gauge_covariance(lib, dev):
    L, N = (4, 4, 4, 4), 3
    S = fields.half_volume(L)
    g = fields.gauge_field(L, N, seed=8)
    om = fields.random_sun((2, S), N, np.random.Generator(np.random.PCG64(9)))
    A = host_complex(gpu_clover(lib, dev, g, L, N, 0.9)[0])
    At = host_complex(gpu_clover(lib, dev, gauge_transform(g, L, om), L, N, 0.9)[0])
    G = np.einsum("ab,psij->psaibj", np.eye(2), om).reshape(2, S, 1, 2 * N, 2 * N)
    want = G @ A @ np.conj(np.swapaxes(G, -1, -2))
    assert np.abs(At - want).max() < 1e-13


@pytest.mark.parametrize("prec", [_lib.DOUBLE, _lib.SINGLE, _lib.HALF, _lib.BF16])
def test_apply(lib, dev, prec):
    L, N, k = (8, 4, 4, 8), 3, 16
    S = fields.half_volume(L)
    g = fields.gauge_field(L, N, seed=10, smear_eps=0.3)
    A = clover_blocks(textbook_clover(g, L, -0.25), N)
    a, ai = gpu_clover(lib, dev, g, L, N, -0.25, prec=prec, inverse=True)
    geom = _lib.geom(L, N)
    psi = fields.gaussian_spinors(L, N, k, seed=11)
    ins = [dev_tensor(psi[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    pin = empty_prec((S, 4, N, k), prec, dev)
    _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), prec, arr, _lib.DOUBLE, k, geom, stream()))
    pout, pback = torch.empty_like(pin), torch.empty_like(pin)
    for parity in (0, 1):
        _lib.check(lib.qcuMrhsCloverApply(pout.data_ptr(), pin.data_ptr(), a.data_ptr(), prec, k, parity, geom,
                                          stream()))
        _lib.check(lib.qcuMrhsCloverApply(pback.data_ptr(), pout.data_ptr(), ai.data_ptr(), prec, k, parity, geom,
                                          stream()))
        outs = [torch.empty((S, 4, N, 2), dtype=torch.float64, device=dev) for _ in range(2 * k)]
        a1, k1 = _lib.ptr_array([t.data_ptr() for t in outs[:k]])
        a2, k2 = _lib.ptr_array([t.data_ptr() for t in outs[k:]])
        _lib.check(lib.qcuMrhsUnpackSpinors(a1, _lib.DOUBLE, pout.data_ptr(), prec, k, geom, stream()))
        _lib.check(lib.qcuMrhsUnpackSpinors(a2, _lib.DOUBLE, pback.data_ptr(), prec, k, geom, stream()))
        torch.cuda.synchronize()
        got = np.stack([host_complex(t) for t in outs[:k]])
        back = np.stack([host_complex(t) for t in outs[k:]])
        v = psi.reshape(k, S, 2, 2 * N)
        want = np.einsum("sbij,ksbj->ksbi", A[parity], v).reshape(psi.shape)
        err = rel_err(got, want)
        err_back = rel_err(back, psi)
        print(f"clover apply prec {prec} parity {parity}: {err:.3e}, A^-1 A psi: {err_back:.3e}")
        assert err < TOL[prec]
        assert err_back < 2 * TOL[prec]


def test_apply_rejects_aliasing(lib, dev):
    L, N = (4, 4, 4, 4), 3
    geom = _lib.geom(L, N)
    buf = torch.zeros(64, device=dev)
    assert lib.qcuMrhsCloverApply(buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), _lib.SINGLE, 1, 0, geom,
                                  stream()) != 0
    assert "alias" in lib.qcuMrhsLastError().decode()


def apply_blocks(blocks, v):
    """blocks [S][2][2N][2N] times spinors [k][S][4][N]."""
    k, S, four, N = v.shape
    w = v.reshape(k, S, 2, 2 * N)
    return np.einsum("sbij,ksbj->ksbi", blocks, w).reshape(v.shape)


def mhat_clover(g, A, Ai, v, L, kappa, dagger):
    """M^ v = A_ee v - kappa^2 D_eo A_oo^-1 D_oe v (dagger: s = -1 hops)."""
    from oracle import oracle

    t = oracle.dslash_many(g, v, L, 1, dagger)
    return apply_blocks(A[0], v) - kappa ** 2 * oracle.dslash_many(g, apply_blocks(Ai[1], t), L, 0, dagger)


@pytest.mark.parametrize("sloppy", [_lib.DOUBLE, _lib.HALF])
def test_clover_cg(lib, dev, sloppy):
    """Even-odd Wilson-clover CG (c_sw = 1, kappa = 0.12): the true residual of
    M^+M^ x = M^+ b, recomputed on the CPU with the textbook clover term and
    the oracle hopping term, is <= 2e-10; mixed and FP64 solutions agree."""
    L, N, k, kappa, csw = (8, 8, 8, 8), 3, 4, 0.12, 1.0
    coeff = -kappa * csw / 2
    g = fields.gauge_field(L, N, seed=21, smear_eps=0.3)
    b = fields.cg_sources(L, N, k, seed=22, n_point=2)
    A = clover_blocks(textbook_clover(g, L, coeff), N)
    Ai = np.linalg.inv(A)
    a, ai = gpu_clover(lib, dev, g, L, N, coeff, inverse=True)
    geom = _lib.geom(L, N)
    gu = dev_tensor(g, dev)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.DOUBLE) // 8, dtype=torch.float64, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.DOUBLE, gu.data_ptr(), _lib.DOUBLE, geom, stream()))
    from mrhs_qcu_amd import fields as F

    bp = dev_tensor(F.to_packed(b), dev)
    sols = {}
    for sl in (_lib.DOUBLE, sloppy):
        xp = torch.empty_like(bp)
        prm = _lib.CgParam()
        prm.kappa, prm.tol, prm.max_iter, prm.reliable_interval, prm.sloppy_prec = kappa, 1e-10, 5000, 50, sl
        res = (ctypes.c_double * k)()
        _lib.check(lib.qcuMrhsCgSolveClover(xp.data_ptr(), bp.data_ptr(), gn.data_ptr(), a.data_ptr(),
                                            ai.data_ptr(), k, geom, ctypes.byref(prm), res, stream()))
        torch.cuda.synchronize()
        assert max(res) <= 1e-10, (sl, list(res))
        x = F.from_packed(host_complex(xp))
        bh = mhat_clover(g, A, Ai, b, L, kappa, 1)
        r = bh - mhat_clover(g, A, Ai, mhat_clover(g, A, Ai, x, L, kappa, 0), L, kappa, 1)
        rel = max(np.linalg.norm(r[i]) / np.linalg.norm(bh[i]) for i in range(k))
        print(f"clover CG sloppy {sl}: {prm.iterations} iterations, {prm.reliable_updates} updates, "
              f"CPU-recomputed residual {rel:.2e}")
        assert rel <= 2e-10
        sols[sl] = x
    for i in range(k):
        d = np.linalg.norm(sols[sloppy][i] - sols[_lib.DOUBLE][i]) / np.linalg.norm(sols[_lib.DOUBLE][i])
        assert d <= 1e-8, (i, d)
