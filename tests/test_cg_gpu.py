"""Multi-RHS even-odd CG on M^+M^ (GPU), checked on the CPU: the FP64 oracle
recomputes ||M^+b - M^+M^ x|| / ||M^+b|| for the returned x.

Tolerance: the solver's own stopping rule is tol = 1e-10 on that relative
residual (recomputed in FP64 on the GPU); the oracle's recomputation must give
<= 2e-10 (it rounds differently), and the mixed-precision solution must agree
with the FP64 solution to <= 1e-8 relative (tol times a condition-number margin).
"""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import dev_tensor, host_complex

pytestmark = pytest.mark.gpu


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def mhat(g, v, L, kappa, dagger):
    """M^ v = v - kappa^2 D_eo D_oe v (dagger: the s = -1 operator)."""
    t = oracle.dslash_many(g, v, L, 1, dagger)
    return v - kappa ** 2 * oracle.dslash_many(g, t, L, 0, dagger)


def solve(lib, dev, g, b, L, N, kappa, sloppy, tol=1e-10, interval=50, delta=0.0, max_iter=5000):
    k = b.shape[0]
    geom = _lib.geom(L, N)
    gu = dev_tensor(g, dev)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.DOUBLE) // 8, dtype=torch.float64, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.DOUBLE, gu.data_ptr(), _lib.DOUBLE, geom, stream()))
    bp = dev_tensor(fields.to_packed(b), dev)
    xp = torch.empty_like(bp)
    prm = _lib.CgParam()
    prm.kappa, prm.tol, prm.max_iter, prm.reliable_interval, prm.sloppy_prec = kappa, tol, max_iter, interval, sloppy
    prm.reliable_delta = delta
    res = (ctypes.c_double * k)()
    _lib.check(lib.qcuMrhsCgSolve(xp.data_ptr(), bp.data_ptr(), gn.data_ptr(), k, geom, ctypes.byref(prm), res,
                                  stream()))
    torch.cuda.synchronize()
    x = fields.from_packed(host_complex(xp))
    return x, prm, list(res)


def check_solution(g, b, x, L, kappa, tol):
    bh = mhat(g, b, L, kappa, 1)
    r = bh - mhat(g, mhat(g, x, L, kappa, 0), L, kappa, 1)
    rel = [np.linalg.norm(r[i]) / np.linalg.norm(bh[i]) for i in range(b.shape[0])]
    return max(rel)


@pytest.mark.parametrize("sloppy", [_lib.DOUBLE, _lib.HALF, _lib.BF16, _lib.SINGLE])
def test_cg_small(lib, dev, sloppy):
    L, N, k, kappa = (8, 8, 8, 8), 3, 6, 0.125
    g = fields.gauge_field(L, N, seed=5)
    b = fields.cg_sources(L, N, k, seed=6, n_point=3)
    x, prm, res = solve(lib, dev, g, b, L, N, kappa, sloppy)
    print(f"sloppy={sloppy}: {prm.iterations} iterations, {prm.reliable_updates} reliable updates, "
          f"true residual max {max(res):.2e}")
    assert max(res) <= 1e-10
    assert check_solution(g, b, x, L, kappa, 1e-10) <= 2e-10


def test_cg_mixed_matches_fp64(lib, dev):
    L, N, k, kappa = (8, 8, 8, 16), 3, 8, 0.135
    g = fields.gauge_field(L, N, seed=7)
    b = fields.cg_sources(L, N, k, seed=8, n_point=4)
    x64, p64, r64 = solve(lib, dev, g, b, L, N, kappa, _lib.DOUBLE)
    x16, p16, r16 = solve(lib, dev, g, b, L, N, kappa, _lib.HALF)
    print(f"fp64 {p64.iterations} it, mixed {p16.iterations} it / {p16.reliable_updates} updates")
    assert max(r64) <= 1e-10 and max(r16) <= 1e-10
    for i in range(k):
        d = np.linalg.norm(x16[i] - x64[i]) / np.linalg.norm(x64[i])
        assert d <= 1e-8, (i, d)


@pytest.mark.parametrize("N", [2, 4])
def test_cg_other_colors(lib, dev, N):
    L, k, kappa = (4, 4, 4, 8), 3, 0.12
    g = fields.gauge_field(L, N, seed=9)
    b = fields.gaussian_spinors(L, N, k, seed=10)
    x, prm, res = solve(lib, dev, g, b, L, N, kappa, _lib.HALF)
    assert max(res) <= 1e-10
    assert check_solution(g, b, x, L, kappa, 1e-10) <= 2e-10


def test_cg_inverter_full_system(lib, dev):
    """qcu.h cg_inverter: (1 - kappa D) x = b on the full lattice."""
    L, N, kappa = (8, 8, 8, 8), 3, 0.125
    S = fields.half_volume(L)
    g = fields.gauge_field(L, N, seed=11)
    b = fields.gaussian_spinors(L, N, 2, seed=12)  # [even, odd]
    gu = dev_tensor(g, dev)
    bd = dev_tensor(b, dev)
    xd = torch.empty_like(bd)
    param = _lib.QcuParam()
    for d in range(4):
        param.lattice_size[d] = L[d]
    grid = _lib.QcuGrid()
    for d in range(4):
        grid.grid_size[d] = 1
    lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
    torch.cuda.synchronize()
    lib.cg_inverter(xd.data_ptr(), bd.data_ptr(), gu.data_ptr(), ctypes.byref(param), 1e-11, kappa)
    torch.cuda.synchronize()
    x = host_complex(xd)
    xe, xo = x[0:1], x[1:2]
    re = xe - kappa * oracle.dslash(g, xo, L, 0, 0) - b[0:1]
    ro = xo - kappa * oracle.dslash(g, xe, L, 1, 0) - b[1:2]
    rel = np.sqrt(np.linalg.norm(re) ** 2 + np.linalg.norm(ro) ** 2) / np.linalg.norm(b)
    print(f"cg_inverter full-system relative residual {rel:.2e}")
    assert rel < 1e-9
    assert x.shape == (2, S, 4, N)


@pytest.mark.parametrize("delta", [3e-3, 1e-3])
def test_cg_half_sparse_reliable_updates_converge(lib, dev, delta):
    """FP16 sloppy CG with rare reliable updates (the iterated residual may
    fall 1e3x between true residuals): the re-orthogonalised search direction
    keeps it converging in about as many iterations as FP64 CG (before that
    step such runs diverged)."""
    L, N, k, kappa = (8, 8, 8, 16), 3, 8, 0.135
    g = fields.gauge_field(L, N, seed=7)
    b = fields.cg_sources(L, N, k, seed=8, n_point=4)
    _, p64, _ = solve(lib, dev, g, b, L, N, kappa, _lib.DOUBLE)
    x, p16, res = solve(lib, dev, g, b, L, N, kappa, _lib.HALF, delta=delta, max_iter=4 * p64.iterations)
    print(f"delta={delta}: fp64 {p64.iterations} it, fp16 {p16.iterations} it / {p16.reliable_updates} updates")
    assert max(res) <= 1e-10
    assert p16.iterations <= 2 * p64.iterations
    assert check_solution(g, b, x, L, kappa, 1e-10) <= 2e-10
