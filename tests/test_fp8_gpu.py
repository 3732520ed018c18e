"""FP8 (OCP E4M3) spinor storage: the gfx950 conversion pinned against the
format's definition, the FP8 operator's error against the FP64 oracle
(stated), and the FP8-sloppy mixed CG reaching the FP64 tolerance.

E4M3 (OCP): 1 sign, 4 exponent bits (bias 7), 3 mantissa bits; subnormals
m/8 * 2^-6; largest finite 448 (code 0x7E); 0x7F / 0xFF are NaN. The library
rounds to nearest even and saturates at +-448.
"""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import dev_tensor, host_complex, rel_err
from test_cg_gpu import check_solution, solve
from test_dslash_gpu import stream

pytestmark = pytest.mark.gpu


def e4m3_values():
    """Value of every code 0..255 (NaN codes as nan)."""
    v = np.empty(256)
    for c in range(256):
        s, e, m = c >> 7, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            v[c] = np.nan
            continue
        mag = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        v[c] = -mag if s else mag
    return v


def e4m3_encode(x):
    """Round to nearest (ties to the even code), saturate at +-448."""
    vals = e4m3_values()
    pos = vals[:127]  # codes 0..126, increasing
    out = np.empty(x.shape, dtype=np.uint8)
    for i, xv in enumerate(x.ravel()):
        a = min(abs(float(xv)), 448.0)
        j = int(np.searchsorted(pos, a))
        if j == 0:
            c = 0
        elif j >= 127:
            c = 126
        else:
            lo, hi = pos[j - 1], pos[j]
            if a - lo < hi - a:
                c = j - 1
            elif a - lo > hi - a:
                c = j
            else:
                c = j - 1 if (j - 1) % 2 == 0 else j
        if np.signbit(xv):  # the sign survives rounding to zero (and -0.0 keeps it)
            c |= 0x80
        out.ravel()[i] = c
    return out


def test_fp8_conversion_matches_format(lib, dev):
    vals = e4m3_values()
    pos = vals[:127]
    mids = (pos[:-1] + pos[1:]) / 2
    rng = np.random.default_rng(3)
    x = np.concatenate([pos, -pos, mids, -mids, np.nextafter(mids, 0), np.nextafter(mids, 1e9),
                        [1e-4, 2.0 ** -10, 2.0 ** -9.5, 500.0, -1e6, 448.0, 464.0], rng.normal(0, 30, 4000),
                        rng.uniform(-0.02, 0.02, 2000)]).astype(np.float32)
    if x.size % 2:
        x = np.append(x, np.float32(0))
    src = torch.from_numpy(x).to(dev)
    dst = torch.zeros(x.size, dtype=torch.uint8, device=dev)
    _lib.check(lib.qcuMrhsConvert(dst.data_ptr(), _lib.FP8, src.data_ptr(), _lib.SINGLE, x.size // 2, stream()))
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    want = e4m3_encode(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:10]]
    # NaN stays NaN (a NaN code, either sign); +-inf saturate to +-448
    sp = torch.tensor([float("nan"), float("inf"), -float("inf"), float("nan")], dtype=torch.float32, device=dev)
    dp = torch.zeros(4, dtype=torch.uint8, device=dev)
    _lib.check(lib.qcuMrhsConvert(dp.data_ptr(), _lib.FP8, sp.data_ptr(), _lib.SINGLE, 2, stream()))
    torch.cuda.synchronize()
    sc = dp.cpu().numpy()
    assert (sc[0] & 0x7F) == 0x7F and (sc[3] & 0x7F) == 0x7F, sc
    assert sc[1] == 0x7E and sc[2] == 0xFE, sc
    # decode every finite code
    codes = np.array([c for c in range(256) if not np.isnan(vals[c])], dtype=np.uint8)
    codes = codes[: codes.size // 2 * 2]
    cd = torch.from_numpy(codes).to(dev)
    back = torch.zeros(codes.size, dtype=torch.float32, device=dev)
    _lib.check(lib.qcuMrhsConvert(back.data_ptr(), _lib.SINGLE, cd.data_ptr(), _lib.FP8, codes.size // 2, stream()))
    torch.cuda.synchronize()
    assert np.array_equal(back.cpu().numpy(), vals[codes].astype(np.float32))


def test_fp8_from_fp64_rounds_once(lib, dev):
    """FP64 -> FP8 is one correct rounding: values a hair (2^-30 relative)
    past every midpoint, which FP32 cannot resolve, still round away from
    the midpoint; exact midpoints tie to the even code."""
    vals = e4m3_values()
    pos = vals[:127]
    mids = (pos[:-1] + pos[1:]) / 2
    eps = 2.0 ** -30
    x = np.concatenate([mids * (1 + eps), mids * (1 - eps), -mids * (1 + eps), mids, pos[1:] * (1 + eps)])
    if x.size % 2:
        x = np.append(x, 0.0)
    src = torch.from_numpy(x.astype(np.float64)).to(dev)
    dst = torch.zeros(x.size, dtype=torch.uint8, device=dev)
    _lib.check(lib.qcuMrhsConvert(dst.data_ptr(), _lib.FP8, src.data_ptr(), _lib.DOUBLE, x.size // 2, stream()))
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    want = e4m3_encode(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:10]]


def test_fp8_dslash_error_stated(lib, dev):
    """16 RHS on eps = 0.3 smeared links at 8^3 x 16, FP8 spinors in and out,
    binary16 links: relative error against the FP64 oracle (stated, < 0.1)."""
    L, N, k = (8, 8, 8, 16), 3, 16
    geom = _lib.geom(L, N)
    g = fields.gauge_field(L, N, seed=21, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, N, k, seed=22)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.FP8) // 2, dtype=torch.float16, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.FP8, dev_tensor(g, dev).data_ptr(), _lib.DOUBLE, geom,
                                    stream()))
    ins = [dev_tensor(psi[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    nbytes = lib.qcuMrhsSpinorBytes(geom, k, _lib.FP8)
    pin = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    pout = torch.empty_like(pin)
    _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), _lib.FP8, arr, _lib.DOUBLE, k, geom, stream()))
    for parity, dagger in ((0, 0), (1, 1)):
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, parity, dagger,
                                     geom, None, stream()))
        outs = [torch.empty(psi.shape[1:] + (2,), dtype=torch.float64, device=dev) for _ in range(k)]
        oarr, okeep = _lib.ptr_array([t.data_ptr() for t in outs])
        _lib.check(lib.qcuMrhsUnpackSpinors(oarr, _lib.DOUBLE, pout.data_ptr(), _lib.FP8, k, geom, stream()))
        torch.cuda.synchronize()
        got = np.stack([host_complex(t) for t in outs])
        err = rel_err(got, oracle.dslash(g, psi, L, parity, dagger))
        print(f"FP8 dslash (binary16 links), 8^3x16 k=16 smeared links, parity={parity}: "
              f"error vs FP64 oracle {err:.3e}")
        assert err < 0.1


def test_fp8_sloppy_cg(lib, dev):
    """Mixed CG with FP8 sloppy fields (binary16 links) to 1e-10: true residual
    recomputed by the oracle, solution against the FP64 solve."""
    L, N, k, kappa = (8, 8, 8, 16), 3, 8, 0.135
    g = fields.gauge_field(L, N, seed=7)
    b = fields.cg_sources(L, N, k, seed=8, n_point=4)
    x64, p64, r64 = solve(lib, dev, g, b, L, N, kappa, _lib.DOUBLE)
    x8, p8, r8 = solve(lib, dev, g, b, L, N, kappa, _lib.FP8, max_iter=20 * p64.iterations)
    print(f"FP64 {p64.iterations} it; FP8 sloppy {p8.iterations} it / {p8.reliable_updates} reliable updates")
    assert max(r8) <= 1e-10
    assert check_solution(g, b, x8, L, kappa, 1e-10) <= 2e-10
    d = max(np.linalg.norm(x8[i] - x64[i]) / np.linalg.norm(x64[i]) for i in range(k))
    assert d <= 1e-8, d


@pytest.mark.parametrize("L,N,k", [((4, 4, 4, 8), 3, 3), ((4, 4, 4, 8), 4, 4)])
def test_fp8_rejects_shapes_without_kernels(lib, dev, L, N, k):
    """FP8 kernels exist for SU(3) with an even RHS count: other shapes fail
    with a message, before any launch."""
    geom = _lib.geom(L, N)
    gn = torch.zeros(lib.qcuMrhsGaugeBytes(geom, _lib.FP8) // 2, dtype=torch.float16, device=dev)
    pin = torch.zeros(lib.qcuMrhsSpinorBytes(geom, k, _lib.FP8), dtype=torch.uint8, device=dev)
    pout = torch.empty_like(pin)
    rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, 0, 0, geom, None, stream())
    assert rc != 0
    assert b"FP8" in lib.qcuMrhsLastError()


def test_fp8_even_rhs_counts(lib, dev):
    """k = 2 (one lane per site) and k = 6 run with the default tile and match
    the FP64 oracle to the FP8 error."""
    L, N = (4, 4, 4, 8), 3
    geom = _lib.geom(L, N)
    g = fields.gauge_field(L, N, seed=31, smear_eps=0.3)
    for k in (2, 6):
        psi = fields.gaussian_spinors(L, N, k, seed=32 + k)
        gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.FP8) // 2, dtype=torch.float16, device=dev)
        _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.FP8, dev_tensor(g, dev).data_ptr(), _lib.DOUBLE, geom,
                                        stream()))
        ins = [dev_tensor(psi[r], dev) for r in range(k)]
        arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
        pin = torch.empty(lib.qcuMrhsSpinorBytes(geom, k, _lib.FP8), dtype=torch.uint8, device=dev)
        pout = torch.empty_like(pin)
        _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), _lib.FP8, arr, _lib.DOUBLE, k, geom, stream()))
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, 1, 0, geom, None,
                                     stream()))
        if k == 2:  # the tuner also finds a tile (one lane per site, links in LDS) and the result stays the same
            t = _lib.Tile()
            _lib.check(lib.qcuMrhsTune(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, geom,
                                       ctypes.byref(t), None, stream()))
            assert t.rhs_per_lane == 2 and t.link_lds == 1
            first = pout.clone()
            _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, 1, 0, geom,
                                         ctypes.byref(t), stream()))
            torch.cuda.synchronize()
            assert torch.equal(first, pout)
        outs = [torch.empty(psi.shape[1:] + (2,), dtype=torch.float64, device=dev) for _ in range(k)]
        oarr, okeep = _lib.ptr_array([t.data_ptr() for t in outs])
        _lib.check(lib.qcuMrhsUnpackSpinors(oarr, _lib.DOUBLE, pout.data_ptr(), _lib.FP8, k, geom, stream()))
        torch.cuda.synchronize()
        err = rel_err(np.stack([host_complex(t) for t in outs]), oracle.dslash(g, psi, L, 1, 0))
        print(f"FP8 k={k}: error vs FP64 oracle {err:.3e}")
        assert err < 0.1
