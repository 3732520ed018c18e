"""Parity at the full sizes of the benchmark configurations, against the FP64
C oracle (oracle/oracle.c, the restatement of the reference's applyWilson,
include/kernels/wilson_dslash_4d.cuh:27-301):

  config 3: 24^3 x 48, SU(3), 16 RHS, links unitarized Gaussian then smeared
            toward the identity (U = V^0.3), FP32 / FP16 / BF16 storage.
            FP32 <= 1e-6; the 16-bit errors are stated (printed) and bounded by
            3e-3 (FP16) and 2e-2 (BF16).
  config 4: even-odd CG on M^+M^ at 32^3 x 64, 16 RHS (12 point sources + 4
            Gaussian), kappa = 0.1350, tol 1e-10: the true residual of each
            solution recomputed by the oracle <= 2e-10, and the mixed-precision
            (FP16 sloppy) solution within 1e-8 of the FP64 one.
  config 5: the uneven batch at its full sizes (SU(2) 16^4 k=8, SU(3)
            24^3x48 k=12, SU(4) 16^3x32 k=24, SU(5) 12^3x24 k=6, FP32): every
            graph replay bit-equal to the same kernels launched eagerly and
            <= 1e-6 from the oracle.

Per-RHS relative error, max over RHS. These tests take about a minute each.
"""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err
from test_cg_gpu import check_solution, solve
from test_dslash_gpu import pack_inputs, stream, unpack

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

C3_L, C3_K = (24, 24, 24, 48), 16
C3_CASES = [(0, 0), (1, 1)]  # (parity, dagger)


@pytest.fixture(scope="module")
def c3():
    """Smeared SU(3) links, Gaussian spinors and the FP64 oracle outputs."""
    print("config 3: generating the smeared 24^3x48 field", flush=True)
    g = fields.gauge_field(C3_L, 3, seed=5, smear_eps=0.3)
    psi = fields.gaussian_spinors(C3_L, 3, C3_K, seed=6)
    want = {pd: oracle.dslash(g, psi, C3_L, pd[0], pd[1]) for pd in C3_CASES}
    return g, psi, want


@pytest.mark.parametrize("prec,bound", [(_lib.SINGLE, 1e-6), (_lib.HALF, 3e-3), (_lib.BF16, 2e-2)])
def test_config3_full_size(lib, dev, c3, prec, bound):
    g, psi, want = c3
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, C3_L, 3, prec)
    pout = torch.empty_like(pin)
    for parity, dagger in C3_CASES:
        print(f"config 3 prec={prec} parity={parity} dagger={dagger}", flush=True)
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, C3_K, parity, dagger,
                                     geom, None, stream()))
        got = unpack(lib, dev, pout, prec, C3_K, geom, psi.shape[1:])
        err = rel_err(got, want[(parity, dagger)])
        print(f"config 3, 24^3x48 k=16 smeared links, prec={prec}: error vs FP64 oracle {err:.3e}", flush=True)
        assert err < bound, err


def test_config4_cg_full_size(lib, dev):
    L, N, k, kappa = (32, 32, 32, 64), 3, 16, 0.1350
    print("config 4: generating 32^3x64 links and sources", flush=True)
    g = fields.gauge_field_device(L, N, seed=0, device=dev).cpu().numpy()
    b = fields.cg_sources(L, N, k, seed=13, n_point=12)
    x64, p64, r64 = solve(lib, dev, g, b, L, N, kappa, _lib.DOUBLE)
    x16, p16, r16 = solve(lib, dev, g, b, L, N, kappa, _lib.HALF)
    print(f"config 4: FP64 {p64.iterations} it, mixed {p16.iterations} it / {p16.reliable_updates} updates; "
          f"GPU true residuals {max(r64):.2e} / {max(r16):.2e}", flush=True)
    assert max(r64) <= 1e-10 and max(r16) <= 1e-10
    for name, x in (("fp64", x64), ("mixed", x16)):
        res = check_solution(g, b, x, L, kappa, 1e-10)
        print(f"config 4 {name}: oracle-recomputed true residual {res:.2e}", flush=True)
        assert res <= 2e-10, (name, res)
    d = max(np.linalg.norm(x16[i] - x64[i]) / np.linalg.norm(x64[i]) for i in range(k))
    print(f"config 4: mixed vs FP64 solution {d:.2e}", flush=True)
    assert d <= 1e-8


C5 = [((16, 16, 16, 16), 2, 8, 0, 0), ((24, 24, 24, 48), 3, 12, 1, 0), ((16, 16, 16, 32), 4, 24, 0, 1),
      ((12, 12, 12, 24), 5, 6, 1, 1)]


def test_config5_batch_full_size(lib, dev):
    items = []
    for i, (L, N, k, parity, dagger) in enumerate(C5):
        print(f"config 5: problem SU({N}) {L} k={k}", flush=True)
        g = fields.gauge_field(L, N, seed=60 + i)
        psi = fields.gaussian_spinors(L, N, k, seed=70 + i)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        want = oracle.dslash(g, psi, L, parity, dagger)
        del g, psi
        items.append(dict(N=N, k=k, parity=parity, dagger=dagger, geom=geom, gnat=gnat, pin=pin,
                          pout=torch.zeros_like(pin), want=want, shape=want.shape[1:]))
    probs = (_lib.Problem * len(items))()
    for p, it in zip(probs, items):
        p.geom = it["geom"]
        p.nrhs, p.parity, p.dagger = it["k"], it["parity"], it["dagger"]
        p.out, p.in_, p.gauge = it["pout"].data_ptr(), it["pin"].data_ptr(), it["gnat"].data_ptr()
    handle = ctypes.c_void_p()
    _lib.check(lib.qcuMrhsBatchCreate(probs, len(items), ctypes.byref(handle), stream()))
    try:
        eager = []
        for i, it in enumerate(items):
            t = _lib.Tile()  # the tile the batch runs this problem with
            _lib.check(lib.qcuMrhsBatchGetTile(handle, i, ctypes.byref(t)))
            o = torch.empty_like(it["pin"])
            _lib.check(lib.qcuMrhsDslash(o.data_ptr(), it["pin"].data_ptr(), it["gnat"].data_ptr(), _lib.SINGLE,
                                         it["k"], it["parity"], it["dagger"], it["geom"], ctypes.byref(t), stream()))
            eager.append(o)
        torch.cuda.synchronize()
        for replay in range(2):
            for it in items:
                it["pout"].fill_(float("nan"))
            _lib.check(lib.qcuMrhsBatchRun(handle, stream()))
            torch.cuda.synchronize()
            for it, o in zip(items, eager):
                assert torch.equal(it["pout"], o), (replay, it["N"])
        for it in items:
            got = unpack(lib, dev, it["pout"], _lib.SINGLE, it["k"], it["geom"], it["shape"])
            err = rel_err(got, it["want"])
            print(f"config 5 SU({it['N']}) k={it['k']}: error vs FP64 oracle {err:.2e}", flush=True)
            assert err < 1e-6, (it["N"], err)
    finally:
        _lib.check(lib.qcuMrhsBatchDestroy(handle))
