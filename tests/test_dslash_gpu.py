"""GPU parity of the dslash paths against the FP64 oracle.

Tolerances (relative error per RHS, max over RHS, vs the FP64 oracle):
FP64 1e-13, FP32 1e-6; FP16 / BF16 storage with FP32 arithmetic are stated
(printed) and bounded loosely (3e-3 / 2e-2).
"""
import ctypes
import math
import itertools

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err, dev_tensor, host_complex, empty_prec, native_gauge

pytestmark = pytest.mark.gpu

TOL = {2: 1e-13, 1: 1e-6, 0: 3e-3, 4: 2e-2}


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def pack_inputs(lib, dev, g_user, psi_user, L, N, prec):
    geom = _lib.geom(L, N)
    gd = dev_tensor(g_user, dev)
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, prec) // _lib.BYTES[prec],), prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), prec, gd.data_ptr(), _lib.DOUBLE, geom, stream()))
    k = psi_user.shape[0]
    ins = [dev_tensor(psi_user[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    pin = empty_prec(psi_user.shape[1:] + (k,), prec, dev)
    _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), prec, arr, _lib.DOUBLE, k, geom, stream()))
    return geom, gnat, pin


def unpack(lib, dev, packed, prec, k, geom, shape):
    outs = [torch.empty(shape + (2,), dtype=torch.float64, device=dev) for _ in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in outs])
    _lib.check(lib.qcuMrhsUnpackSpinors(arr, _lib.DOUBLE, packed.data_ptr(), prec, k, geom, stream()))
    torch.cuda.synchronize()
    return np.stack([host_complex(t) for t in outs])


def run_packed(lib, dev, g, psi, L, N, parity, dagger, prec, tile=None, grid=(1, 1, 1, 1)):
    k = psi.shape[0]
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
    for d in range(4):
        geom.grid[d] = grid[d]
    pout = torch.empty_like(pin)
    _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger, geom,
                                 tile, stream()))
    return unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])


def test_config1_fp64_8x8x8x8(lib, dev):
    """8^4, SU(3), nrhs=1, FP64, both parities and dagger flags: <= 1e-13."""
    L, N = (8, 8, 8, 8), 3
    g = fields.gauge_field(L, N, seed=0)
    psi = fields.gaussian_spinors(L, N, 1, seed=0)
    for parity in (0, 1):
        for dagger in (0, 1):
            got = run_packed(lib, dev, g, psi, L, N, parity, dagger, _lib.DOUBLE)
            want = oracle.dslash(g, psi, L, parity, dagger)
            err = rel_err(got, want)
            print(f"c1 fp64 parity={parity} dagger={dagger} rel_err={err:.3e}")
            assert err < TOL[2]


def test_config2_fp32_16x16x16x32_k12(lib, dev):
    """16^3x32, SU(3), nrhs=12, FP32 vs FP64 oracle: <= 1e-6, both parities."""
    L, N, k = (16, 16, 16, 32), 3, 12
    g = fields.gauge_field(L, N, seed=0)
    psi = fields.gaussian_spinors(L, N, k, seed=0)
    for parity, dagger in ((0, 0), (1, 0), (0, 1)):
        got = run_packed(lib, dev, g, psi, L, N, parity, dagger, _lib.SINGLE)
        want = oracle.dslash(g, psi, L, parity, dagger)
        err = rel_err(got, want)
        print(f"c2 fp32 parity={parity} dagger={dagger} rel_err={err:.3e}")
        assert err < TOL[1]


@pytest.mark.parametrize("prec", [_lib.HALF, _lib.BF16])
def test_low_precision_storage_error_stated(lib, dev, prec):
    """FP16 / BF16 storage, FP32 arithmetic, smeared links (eps = 0.3)."""
    L, N, k = (8, 8, 8, 16), 3, 16
    g = fields.gauge_field(L, N, seed=0, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, N, k, seed=0)
    got = run_packed(lib, dev, g, psi, L, N, 0, 0, prec)
    want = oracle.dslash(g, psi, L, 0, 0)
    err = rel_err(got, want)
    print(f"{'fp16' if prec == 0 else 'bf16'} storage rel_err vs fp64 oracle = {err:.3e}")
    assert err < TOL[prec]


@pytest.mark.parametrize("N", [2, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("prec", [_lib.DOUBLE, _lib.SINGLE])
def test_colors(lib, dev, N, prec):
    L = (4, 4, 4, 8)
    k = 4
    g = fields.gauge_field(L, N, seed=N)
    psi = fields.gaussian_spinors(L, N, k, seed=N + 1)
    for parity in (0, 1):
        got = run_packed(lib, dev, g, psi, L, N, parity, 1, prec)
        want = oracle.dslash(g, psi, L, parity, 1)
        assert rel_err(got, want) < TOL[prec]


def test_every_tile_agrees(lib, dev):
    """Every instantiated tile (R, LDS staging, block, t/z blocking, both
    traversal orders) gives the oracle's result (FP32, N = 3, k = 8 and 6)."""
    L, N = (8, 4, 6, 8), 3
    g = fields.gauge_field(L, N, seed=3)
    for k in (8, 6):
        psi = fields.gaussian_spinors(L, N, k, seed=4)
        want = oracle.dslash(g, psi, L, 1, 0)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        n = 0
        for R in (1, 2, 4):
            for lds in (0, 1):
                for block in (64, 128, 192, 256):
                    for (tb, zb), trav in itertools.product(((1, 0), (2, 0), (4, 0), (2, 1), (4, 2), (2, 3)), (0, 1)):
                        if k % R or L[3] % tb:
                            continue
                        hints = (block // 64) & 3
                        pout = torch.zeros_like(pin)
                        rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, 1, 0,
                                               geom, _lib.tile(R, block, lds, tb, zb, hints, 0, trav), stream())
                        if rc != 0:
                            msg = lib.qcuMrhsLastError().decode()
                            assert "no kernel instantiated" in msg or "LDS" in msg, msg
                            continue
                        got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
                        assert rel_err(got, want) < TOL[1], (R, lds, block, tb, zb, hints, trav)
                        n += 1
        assert n >= 96


@pytest.mark.parametrize("prec", [_lib.HALF, _lib.BF16])
def test_low_precision_every_tile(lib, dev, prec):
    """FP16 (packed-half dot-product path) and BF16 storage: every
    instantiated tile, N = 2 and 3, both parities and dagger signs, within the
    stated storage tolerance of the FP64 oracle; all tiles agree with each
    other to a few half-precision roundings."""
    for N, L, k in ((3, (8, 4, 6, 8), 8), (2, (4, 4, 4, 8), 4)):
        g = fields.gauge_field(L, N, seed=11 + N, smear_eps=0.3)
        psi = fields.gaussian_spinors(L, N, k, seed=12 + N)
        for parity, dagger in ((0, 0), (1, 1)):
            want = oracle.dslash(g, psi, L, parity, dagger)
            geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
            n = 0
            first = None
            for R, lds, block, (tb, zb), trav in itertools.product((1, 2, 4), (0, 1), (64, 256), ((1, 0), (2, 3)),
                                                                    (0, 1)):
                if k % R or L[3] % tb or (zb and L[2] % zb):
                    continue
                pout = torch.zeros_like(pin)
                rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger,
                                       geom, _lib.tile(R, block, lds, tb, zb, 0, 0, trav), stream())
                if rc != 0:
                    msg = lib.qcuMrhsLastError().decode()
                    assert "no kernel instantiated" in msg or "LDS" in msg, msg
                    continue
                got = unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])
                assert rel_err(got, want) < TOL[prec], (N, R, lds, block, tb, zb, trav)
                if first is None:
                    first = got
                else:
                    assert rel_err(got, first) < TOL[prec] / 4, (N, R, lds, block, tb, zb, trav)
                n += 1
            assert n >= 8, (N, n)


@pytest.mark.parametrize("grid", [(2, 1, 1, 1), (1, 1, 1, 2), (2, 2, 2, 2)])
def test_boundary_skip(lib, dev, grid):
    L, N, k = (4, 4, 4, 4), 3, 2
    g = fields.gauge_field(L, N, seed=9)
    psi = fields.gaussian_spinors(L, N, k, seed=10)
    for parity in (0, 1):
        got = run_packed(lib, dev, g, psi, L, N, parity, 0, _lib.DOUBLE, grid=grid)
        want = oracle.dslash(g, psi, L, parity, 0, grid=grid)
        assert rel_err(got, want) < TOL[2]


def test_xpay(lib, dev):
    L, N, k = (4, 4, 4, 8), 3, 4
    g = fields.gauge_field(L, N, seed=11)
    psi = fields.gaussian_spinors(L, N, k, seed=12)
    x = fields.gaussian_spinors(L, N, k, seed=13)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.DOUBLE)
    _, _, px = pack_inputs(lib, dev, g, x, L, N, _lib.DOUBLE)
    pout = torch.empty_like(pin)
    a, b = 1.0, -0.135 ** 2
    _lib.check(lib.qcuMrhsDslashXpay(pout.data_ptr(), pin.data_ptr(), px.data_ptr(), a, b, gnat.data_ptr(),
                                     _lib.DOUBLE, k, 0, 1, geom, None, stream()))
    got = unpack(lib, dev, pout, _lib.DOUBLE, k, geom, psi.shape[1:])
    want = a * x + b * oracle.dslash(g, psi, L, 0, 1)
    assert rel_err(got, want) < TOL[2]


def test_pack_roundtrip_and_gauge_layout(lib, dev):
    L, N, k = (4, 4, 6, 4), 3, 5
    g = fields.gauge_field(L, N, seed=14)
    psi = fields.gaussian_spinors(L, N, k, seed=15)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.DOUBLE)
    # packed layout [S][4][N][k], bit-exact
    assert np.array_equal(host_complex(pin), fields.to_packed(psi))
    # native gauge layout, bit-exact
    assert np.array_equal(host_complex(gnat).reshape(2, -1, 8, N, N), native_gauge(g, L))
    # round trip, bit-exact
    back = unpack(lib, dev, pin, _lib.DOUBLE, k, geom, psi.shape[1:])
    assert np.array_equal(back, psi)
    # FP32 / FP16 packing rounds to nearest even (numpy's casts)
    for prec, npt in ((_lib.SINGLE, np.float32), (_lib.HALF, np.float16)):
        _, _, p = pack_inputs(lib, dev, g, psi, L, N, prec)
        host = p.detach().cpu().numpy()
        want = fields.to_packed(psi)
        assert np.array_equal(host[..., 0], want.real.astype(npt))
        assert np.array_equal(host[..., 1], want.imag.astype(npt))
    # BF16 packing from FP64 is one correct rounding (nearest, ties to even)
    _, _, p = pack_inputs(lib, dev, g, psi, L, N, _lib.BF16)
    bits = p.detach().view(torch.int16).cpu().numpy().view(np.uint16)
    want = fields.to_packed(psi)
    assert np.array_equal(bits[..., 0], bf16_bits(want.real))
    assert np.array_equal(bits[..., 1], bf16_bits(want.imag))


@pytest.mark.parametrize("k", [1, 2, 16, 32, 5, 33, 48])
def test_pack_unpack_every_path(lib, dev, k):
    """Packing through the LDS tile (k a power of two <= 32; 2x2x2x6 has 24
    sites x 12 rows = 288 rows, so the last 64-row tile is partial), by rows
    (other k) and in chunks of 32 RHS (k > 32): packed layout and round trip
    bit-exact in FP64, FP32 packing from FP64 equal to numpy's cast."""
    L, N = (2, 2, 2, 6), 3
    psi = fields.gaussian_spinors(L, N, k, seed=100 + k)
    geom = _lib.geom(L, N)
    ins = [dev_tensor(psi[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    want = fields.to_packed(psi)
    for prec, npt in ((_lib.DOUBLE, np.float64), (_lib.SINGLE, np.float32)):
        pin = empty_prec(psi.shape[1:] + (k,), prec, dev)
        _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), prec, arr, _lib.DOUBLE, k, geom, stream()))
        torch.cuda.synchronize()
        host = pin.detach().cpu().numpy()
        assert np.array_equal(host[..., 0], want.real.astype(npt)), prec
        assert np.array_equal(host[..., 1], want.imag.astype(npt)), prec
        back = unpack(lib, dev, pin, prec, k, geom, psi.shape[1:])
        assert np.array_equal(back, psi.astype(np.complex64) if prec == _lib.SINGLE else psi), prec


def bf16_round(x):
    """Correctly rounded bfloat16 value of a double (8 significant bits,
    nearest, ties to even; the normal range of FP32), by exact scaling."""
    if x == 0 or not math.isfinite(x):
        return x
    m, e = math.frexp(x)  # x = m 2^e, 0.5 <= |m| < 1
    ulp = math.ldexp(1.0, e - 8)
    return round(x / ulp) * ulp  # x / ulp is exact; round() ties to even


def bf16_bits(a):
    a = np.asarray(a, dtype=np.float64)
    r = np.array([bf16_round(float(v)) for v in a.ravel()]).reshape(a.shape)
    return (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_bf16_from_fp64_rounds_once(lib, dev):
    """FP64 -> BF16: values a hair (2^-30 relative) past a bfloat16 midpoint,
    which FP32 cannot tell from the midpoint, still round away from it;
    exact midpoints tie to even."""
    j = np.arange(0, 128)
    mids = np.concatenate([(1 + (j + 0.5) / 128) * s for s in (1.0, 2.0 ** -9, 2.0 ** 13)])
    eps = 2.0 ** -30
    x = np.concatenate([mids * (1 + eps), mids * (1 - eps), -mids * (1 + eps), mids])
    src = torch.from_numpy(x).to(dev)
    dst = torch.zeros(x.size, dtype=torch.int16, device=dev)
    _lib.check(lib.qcuMrhsConvert(dst.data_ptr(), _lib.BF16, src.data_ptr(), _lib.DOUBLE, x.size // 2, stream()))
    torch.cuda.synchronize()
    got = dst.cpu().numpy().view(np.uint16)
    want = bf16_bits(x)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:10]]


def test_user_layout_path(lib, dev):
    """qcuMrhsDslashUser: user AoS spinors in place, native gauge."""
    L, N, k = (8, 4, 4, 8), 3, 5
    g = fields.gauge_field(L, N, seed=16)
    psi = fields.gaussian_spinors(L, N, k, seed=17)
    for prec in (_lib.DOUBLE, _lib.SINGLE):
        geom, gnat, _ = pack_inputs(lib, dev, g, psi[:1], L, N, prec)
        cdt = np.complex128 if prec == _lib.DOUBLE else np.complex64
        ins = [dev_tensor(psi[r].astype(cdt), dev) for r in range(k)]
        outs = [torch.empty_like(t) for t in ins]
        ai, k1 = _lib.ptr_array([t.data_ptr() for t in ins])
        ao, k2 = _lib.ptr_array([t.data_ptr() for t in outs])
        _lib.check(lib.qcuMrhsDslashUser(ao, ai, gnat.data_ptr(), prec, k, 1, 0, geom, stream()))
        torch.cuda.synchronize()
        got = np.stack([host_complex(t) for t in outs])
        assert rel_err(got, oracle.dslash(g, psi, L, 1, 0)) < TOL[prec]


def test_reference_layout_path(lib, dev):
    """qcu::instantiateDslash on reference SoA spinors and SoA gauge."""
    L, N, k = (4, 4, 4, 8), 4, 3
    g = fields.gauge_field(L, N, seed=18)
    psi = fields.gaussian_spinors(L, N, k, seed=19)
    S = fields.half_volume(L)
    gsoa = oracle.shift_gauge(g.reshape(-1), True, S, N)
    gd = dev_tensor(gsoa, dev)
    ins = [dev_tensor(oracle.shift_vector(psi[r].reshape(-1), True, S, N), dev) for r in range(k)]
    outs = [torch.empty_like(t) for t in ins]
    ai, k1 = _lib.ptr_array([t.data_ptr() for t in ins])
    ao, k2 = _lib.ptr_array([t.data_ptr() for t in outs])
    geom = _lib.geom(L, N)
    _lib.check(lib.qcuMrhsDslashReferenceLayout(ao, ai, gd.data_ptr(), _lib.DOUBLE, k, 0, 1, geom))
    got = np.stack([oracle.shift_vector(host_complex(t), False, S, N).reshape(S, 4, N) for t in outs])
    assert rel_err(got, oracle.dslash(g, psi, L, 0, 1)) < TOL[2]


def test_shift_kernels_bit_exact(lib, dev):
    L, N = (6, 4, 4, 4), 3
    S = fields.half_volume(L)
    v = fields.gaussian_spinors(L, N, 1, seed=20)[0].reshape(-1)
    g = fields.gauge_field(L, N, seed=21).reshape(-1)
    geom = _lib.geom(L, N)
    vd, gd = dev_tensor(v, dev), dev_tensor(g, dev)
    vs, gs = torch.empty_like(vd), torch.empty_like(gd)
    _lib.check(lib.qcuMrhsShiftSpinor(vs.data_ptr(), vd.data_ptr(), 0, geom, 16, stream()))
    _lib.check(lib.qcuMrhsShiftGauge(gs.data_ptr(), gd.data_ptr(), 0, geom, 16, stream()))
    torch.cuda.synchronize()
    assert np.array_equal(host_complex(vs), oracle.shift_vector(v, True, S, N))
    assert np.array_equal(host_complex(gs), oracle.shift_gauge(g, True, S, N))
    vb, gb = torch.empty_like(vd), torch.empty_like(gd)
    _lib.check(lib.qcuMrhsShiftSpinor(vb.data_ptr(), vs.data_ptr(), 1, geom, 16, stream()))
    _lib.check(lib.qcuMrhsShiftGauge(gb.data_ptr(), gs.data_ptr(), 1, geom, 16, stream()))
    torch.cuda.synchronize()
    assert torch.equal(vb, vd) and torch.equal(gb, gd)


def test_qcu_h_call_sequence(lib, dev):
    """The reference's src/test.cu:39-48 sequence with seeded data, k = 3,
    FP64 and FP32 buffers: initGridSize -> loadQcuGauge -> getDslash ->
    pushBackFermions x k -> startDslash."""
    L, N, k = (8, 8, 8, 8), 3, 3
    g = fields.gauge_field(L, N, seed=0)
    psi = fields.gaussian_spinors(L, N, k, seed=22)
    grid = _lib.QcuGrid()
    param = _lib.QcuParam()
    for d in range(4):
        grid.grid_size[d] = 1
        param.lattice_size[d] = L[d]
    lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
    gd = dev_tensor(g, dev)
    lib.loadQcuGauge(gd.data_ptr())
    for prec, cdt in ((2, np.complex128), (1, np.complex64)):
        lib.getDslash(0, -3.5, N, k, prec)
        ins = [dev_tensor(psi[r].astype(cdt), dev) for r in range(k)]
        outs = [torch.full_like(t, float("nan")) for t in ins]
        torch.cuda.synchronize()
        for r in range(k):
            lib.pushBackFermions(outs[r].data_ptr(), ins[r].data_ptr())
        # inputs are consumed at push time
        for t in ins:
            t.zero_()
        torch.cuda.synchronize()
        for parity in (0,):
            lib.startDslash(parity, 0)
            got = np.stack([host_complex(t) for t in outs])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, 0)) < TOL[prec]


def test_qcu_h_many_rhs_chunks(lib, dev):
    """startDslash with more queued right-hand sides than one pointer pack
    holds (40 > 32: two launches), N = 2 and N = 4, against the oracle; the
    queue is empty afterwards and a second round works."""
    for N, L, k in ((2, (4, 4, 4, 4), 40), (4, (4, 4, 4, 8), 35)):
        g = fields.gauge_field(L, N, seed=30 + N)
        psi = fields.gaussian_spinors(L, N, k, seed=31 + N)
        grid = _lib.QcuGrid()
        param = _lib.QcuParam()
        for d in range(4):
            grid.grid_size[d] = 1
            param.lattice_size[d] = L[d]
        lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
        gd = dev_tensor(g, dev)
        lib.loadQcuGauge(gd.data_ptr())
        lib.getDslash(0, 0.0, N, k, 2)
        ins = [dev_tensor(psi[r], dev) for r in range(k)]
        for parity, dagger in ((1, 0), (0, 1)):
            outs = [torch.full_like(t, float("nan")) for t in ins]
            torch.cuda.synchronize()
            for r in range(k):
                lib.pushBackFermions(outs[r].data_ptr(), ins[r].data_ptr())
            lib.startDslash(parity, dagger)
            got = np.stack([host_complex(t) for t in outs])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger)) < TOL[2], (N, parity, dagger)


def test_qcu_h_push_buffer_reused_at_once(lib, dev):
    """pushBackFermions returns only once the input has been read: a caller
    that refills one input buffer for the next RHS right after each push, on
    a stream that does not wait for the library's, still gets every RHS's own
    result. 24^3x48 FP64 (64 MB per RHS, so the push copy takes ~10 us),
    k = 3, against the oracle."""
    L, N, k = (24, 24, 24, 48), 3, 3
    g = fields.gauge_field(L, N, seed=41)
    psi = fields.gaussian_spinors(L, N, k, seed=42)
    grid = _lib.QcuGrid()
    param = _lib.QcuParam()
    for d in range(4):
        grid.grid_size[d] = 1
        param.lattice_size[d] = L[d]
    lib.initGridSize(ctypes.byref(grid), ctypes.byref(param), N)
    gd = dev_tensor(g, dev)
    lib.loadQcuGauge(gd.data_ptr())
    lib.getDslash(0, 0.0, N, k, 2)
    srcs = [dev_tensor(psi[r], dev) for r in range(k)]
    buf = torch.empty_like(srcs[0])
    outs = [torch.full_like(t, float("nan")) for t in srcs]
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    for r in range(k):
        with torch.cuda.stream(side):
            buf.copy_(srcs[r])
        side.synchronize()
        lib.pushBackFermions(outs[r].data_ptr(), buf.data_ptr())
    with torch.cuda.stream(side):
        buf.fill_(float("nan"))
    side.synchronize()
    lib.startDslash(1, 0)
    got = np.stack([host_complex(t) for t in outs])
    assert rel_err(got, oracle.dslash(g, psi, L, 1, 0)) < TOL[2]


@pytest.mark.parametrize("prec", [_lib.SINGLE, _lib.HALF])
def test_full_size_properties(lib, dev, prec):
    """24^3x48, nrhs = 16: RHS independence (a permutation of the RHS permutes
    the output) and exact homogeneity (doubling the input doubles the output
    bit for bit)."""
    L, N, k = (24, 24, 24, 48), 3, 16
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    torch.manual_seed(0)
    gu = torch.randn(4, 2, S, N, N, 2, dtype=torch.float64, device=dev)
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, prec) // _lib.BYTES[prec],), prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), prec, gu.data_ptr(), _lib.DOUBLE, geom, stream()))
    pin = torch.randn(S, 4, N, k, 2, device=dev).to(gnat.dtype)
    out = torch.empty_like(pin)
    _lib.check(lib.qcuMrhsDslash(out.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, 0, 0, geom, None,
                                 stream()))
    perm = torch.randperm(k, device=dev)
    pin2 = pin[:, :, :, perm].contiguous()
    out2 = torch.empty_like(pin2)
    _lib.check(lib.qcuMrhsDslash(out2.data_ptr(), pin2.data_ptr(), gnat.data_ptr(), prec, k, 0, 0, geom, None,
                                 stream()))
    torch.cuda.synchronize()
    # Each RHS is computed independently; two RHS that share a lane may round
    # differently where the compiler contracts a multiply-add into an FMA for
    # one slot and not the other, so the check is to one storage ulp.
    a = out2.double().permute(3, 0, 1, 2, 4).reshape(k, -1)
    b = out[:, :, :, perm].double().permute(3, 0, 1, 2, 4).reshape(k, -1)
    err = float(((a - b).norm(dim=1) / b.norm(dim=1)).max())
    ulp = 2.0 ** -11 if prec == _lib.HALF else 2.0 ** -24
    assert err <= ulp, f"RHS permutation changed the result by {err:.3e} (unit roundoff {ulp:.1e})"
    if prec == _lib.SINGLE:
        pin3 = (2.0 * pin).contiguous()
        out3 = torch.empty_like(pin3)
        _lib.check(lib.qcuMrhsDslash(out3.data_ptr(), pin3.data_ptr(), gnat.data_ptr(), prec, k, 0, 0, geom, None,
                                     stream()))
        torch.cuda.synchronize()
        # scaling by 2 is exact in binary floating point, FMA or not
        assert bool(torch.equal(out3, 2.0 * out)), "2x input did not give exactly 2x output"


@pytest.mark.parametrize("in_prec,out_prec", [(_lib.HALF, _lib.HALF), (_lib.HALF, _lib.SINGLE),
                                              (_lib.BF16, _lib.SINGLE), (_lib.SINGLE, _lib.SINGLE),
                                              (_lib.DOUBLE, _lib.DOUBLE)])
def test_dslash_ex_scales_xpay_norms(lib, dev, in_prec, out_prec):
    """qcuMrhsDslashEx: out = a x + b D in with per-RHS storage scales and
    per-RHS |out|^2, against the FP64 oracle."""
    L, N, k = (4, 4, 4, 8), 3, 4
    g = fields.gauge_field(L, N, seed=23)
    psi = fields.gaussian_spinors(L, N, k, seed=24)
    x = fields.gaussian_spinors(L, N, k, seed=25)
    a, b = 0.75, -0.135 ** 2
    sc_in = np.array([0.5, 2.0, 0.25, 4.0], dtype=np.float32)
    sc_x = np.array([1.5, 0.125, 8.0, 1.0], dtype=np.float32)
    sc_out = np.array([16.0, 0.5, 3.0, 0.0625], dtype=np.float32)
    scaled = in_prec == _lib.HALF
    geom = _lib.geom(L, N)
    gd = dev_tensor(g, dev)
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, in_prec) // _lib.BYTES[in_prec],), in_prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), in_prec, gd.data_ptr(), _lib.DOUBLE, geom, stream()))
    si = sc_in if scaled else np.ones(k, np.float32)
    sx = sc_x if scaled else np.ones(k, np.float32)
    so = sc_out if (scaled and out_prec == _lib.HALF) else np.ones(k, np.float32)
    pin_h = fields.to_packed(psi / si[:, None, None, None])
    px_h = fields.to_packed(x / sx[:, None, None, None])
    conv = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}
    pin = dev_tensor(pin_h, dev).to(conv[in_prec])
    px = dev_tensor(px_h, dev).to(conv[in_prec])
    pout = empty_prec(pin_h.shape, out_prec, dev)
    dsi, dsx, dso = (torch.from_numpy(v).to(dev) for v in (si, sx, so))
    nrm = torch.zeros(k, dtype=torch.float64, device=dev)
    _lib.check(lib.qcuMrhsDslashEx(pout.data_ptr(), out_prec, pin.data_ptr(), px.data_ptr(), in_prec, a, b,
                                   dsi.data_ptr(), dsx.data_ptr(), dso.data_ptr(), nrm.data_ptr(), gnat.data_ptr(),
                                   k, 0, 1, geom, stream()))
    torch.cuda.synchronize()
    got = fields.from_packed(host_complex(pout)) * so[:, None, None, None]
    # expected, with the inputs as rounded to their storage precision
    psi_r = fields.from_packed(host_complex(pin)) * si[:, None, None, None]
    x_r = fields.from_packed(host_complex(px)) * sx[:, None, None, None]
    want = a * x_r + b * oracle.dslash(g, psi_r, L, 0, 1)
    tol = {_lib.HALF: 3e-3, _lib.BF16: 2e-2, _lib.SINGLE: 1e-6, _lib.DOUBLE: 1e-13}
    err = rel_err(got, want)
    print(f"in={in_prec} out={out_prec} rel_err={err:.2e}")
    assert err < max(tol[in_prec], tol[out_prec])
    wn = np.array([np.vdot(want[r], want[r]).real for r in range(k)])
    assert np.allclose(nrm.cpu().numpy(), wn, rtol=max(tol[in_prec], tol[out_prec]) * 4)


@pytest.mark.parametrize("L", [(2, 2, 2, 2), (2, 4, 2, 6), (6, 2, 4, 2), (10, 6, 2, 4)])
def test_extreme_lattice_shapes(lib, dev, L):
    """Smallest and ragged extents (x half-extent 1, 3, 5; extents of 2 in
    y, z or t, where both neighbours in a direction are the same site) with
    the default tile and with every dimension split (boundary skip): FP64
    against the oracle, both parities and dagger signs."""
    N, k = 3, 3
    g = fields.gauge_field(L, N, seed=31)
    psi = fields.gaussian_spinors(L, N, k, seed=32)
    for grid in ((1, 1, 1, 1), (2, 2, 2, 2)):
        for parity, dagger in ((0, 0), (1, 1)):
            got = run_packed(lib, dev, g, psi, L, N, parity, dagger, _lib.DOUBLE, grid=grid)
            want = oracle.dslash(g, psi, L, parity, dagger, grid=grid)
            assert rel_err(got, want) < TOL[2], (L, grid, parity, dagger)


@pytest.mark.parametrize("N", [5, 6, 8])
@pytest.mark.parametrize("prec", [_lib.DOUBLE, _lib.SINGLE, _lib.HALF, _lib.BF16])
def test_large_n_link_rows(lib, dev, N, prec):
    """N >= 5 reads each link row by row inside the mat-vec, from LDS or from
    global memory: both tile kinds (and R = 2 where instantiated) against the
    oracle."""
    # k = 16 and 64-lane workgroups keep the LDS-staged links of a workgroup
    # (a few sites of 8 N^2 complex) within the per-workgroup LDS budget.
    L, k = (4, 4, 4, 8), 16
    g = fields.gauge_field(L, N, seed=40 + N, smear_eps=0.3 if prec in (_lib.HALF, _lib.BF16) else None)
    psi = fields.gaussian_spinors(L, N, k, seed=41 + N)
    want = oracle.dslash(g, psi, L, 1, 0)
    tiles = [_lib.tile(R=1, block=64, lds=0), _lib.tile(R=1, block=64, lds=1)]
    if N == 5:
        tiles.append(_lib.tile(R=2, block=64, lds=1))
    for t in tiles:
        got = run_packed(lib, dev, g, psi, L, N, 1, 0, prec, tile=t)
        err = rel_err(got, want)
        assert err < TOL[prec], (t.as_tuple(), err)


def test_xcd_cyclic_order(lib, dev):
    """Cache hint bit 3 (XCDs side by side on adjacent halves of y-streamed
    blocks) only reorders workgroups: the output equals the oracle at a small
    size, and equals the XCD-contiguous order bit for bit at 24^3x48,
    nrhs = 16, FP32 with the tiles the tuner offers there. A tile the lattice
    does not split into whole chunks runs in the contiguous order."""
    L, N, k = (8, 4, 6, 8), 3, 8
    g = fields.gauge_field(L, N, seed=21)
    psi = fields.gaussian_spinors(L, N, k, seed=22)
    want = oracle.dslash(g, psi, L, 0, 1)
    # (R, block, lds, t_block, z_block): (2, 64, 1, 2, 3) splits into chunks of 3 workgroups;
    # (2, 256, 1, 2, 3) does not split (half a block is 192 lanes).
    for R, block, lds, tb, zb in ((2, 64, 1, 2, 3), (2, 64, 0, 2, 3), (1, 128, 1, 2, 2), (2, 256, 1, 2, 3)):
        got = run_packed(lib, dev, g, psi, L, N, 0, 1, _lib.SINGLE, tile=_lib.tile(R, block, lds, tb, zb, 8, 0, 1))
        assert rel_err(got, want) < TOL[1], (R, block, lds, tb, zb)
    for prec, R in ((_lib.DOUBLE, 1), (_lib.DOUBLE, 2), (_lib.HALF, 2), (_lib.BF16, 2)):
        got = run_packed(lib, dev, g, psi, L, N, 0, 1, prec, tile=_lib.tile(R, 64, 1, 2, 3, 8, 0, 1))
        assert rel_err(got, want) < TOL[prec], (prec, R)

    L, k = (24, 24, 24, 48), 16
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    torch.manual_seed(1)
    gu = torch.randn(4, 2, S, N, N, 2, dtype=torch.float64, device=dev)
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, _lib.SINGLE) // 4,), _lib.SINGLE, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), _lib.SINGLE, gu.data_ptr(), _lib.DOUBLE, geom, stream()))
    pin = torch.randn(S, 4, N, k, 2, device=dev)
    for tb, zb, block in ((6, 6, 128), (3, 6, 128), (4, 6, 256), (2, 6, 128)):
        outs = []
        for hints in (0, 8):
            out = torch.full_like(pin, float("nan"))
            _lib.check(lib.qcuMrhsDslash(out.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, 1, 0, geom,
                                         _lib.tile(2, block, 1, tb, zb, hints, 0, 1), stream()))
            outs.append(out)
        torch.cuda.synchronize()
        assert bool(torch.equal(outs[0], outs[1])), (tb, zb, block)


@pytest.mark.parametrize("in_prec,out_prec", [(_lib.HALF, _lib.HALF), (_lib.HALF, _lib.SINGLE),
                                              (_lib.BF16, _lib.BF16), (_lib.BF16, _lib.SINGLE),
                                              (_lib.SINGLE, _lib.SINGLE), (_lib.DOUBLE, _lib.DOUBLE)])
def test_dslash_ex_scaled_without_xpay(lib, dev, in_prec, out_prec):
    """qcuMrhsDslashEx with x = NULL: stored out = b D(sc_in * in) / sc_out and
    |b D(sc_in * in)|^2 per RHS, at 16 RHS (the fixed-RHS-count kernels; for
    16-bit input the default tile stages the per-RHS epilogue factors in LDS
    beside the links), against the FP64 oracle."""
    L, N, k = (4, 4, 4, 8), 3, 16
    g = fields.gauge_field(L, N, seed=33)
    psi = fields.gaussian_spinors(L, N, k, seed=34)
    b = -0.135 ** 2
    rng = np.random.default_rng(35)
    si = (2.0 ** rng.integers(-3, 4, size=k)).astype(np.float32)
    so = (2.0 ** rng.integers(-4, 3, size=k) * 0.75).astype(np.float32)
    geom = _lib.geom(L, N)
    gd = dev_tensor(g, dev)
    gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, in_prec) // _lib.BYTES[in_prec],), in_prec, dev)
    _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), in_prec, gd.data_ptr(), _lib.DOUBLE, geom, stream()))
    conv = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}
    pin = dev_tensor(fields.to_packed(psi / si[:, None, None, None]), dev).to(conv[in_prec])
    pout = empty_prec(pin.shape[:-1], out_prec, dev)
    dsi, dso = (torch.from_numpy(v).to(dev) for v in (si, so))
    nrm = torch.zeros(k, dtype=torch.float64, device=dev)
    for parity, dagger in ((0, 1), (1, 0)):
        _lib.check(lib.qcuMrhsDslashEx(pout.data_ptr(), out_prec, pin.data_ptr(), None, in_prec, 0.0, b,
                                       dsi.data_ptr(), None, dso.data_ptr(), nrm.data_ptr(), gnat.data_ptr(), k,
                                       parity, dagger, geom, stream()))
        torch.cuda.synchronize()
        got = fields.from_packed(host_complex(pout)) * so[:, None, None, None]
        psi_r = fields.from_packed(host_complex(pin)) * si[:, None, None, None]
        want = b * oracle.dslash(g, psi_r, L, parity, dagger)
        tol = {_lib.HALF: 3e-3, _lib.BF16: 2e-2, _lib.SINGLE: 1e-6, _lib.DOUBLE: 1e-13}
        t = max(tol[in_prec], tol[out_prec])
        err = rel_err(got, want)
        wn = np.array([np.vdot(want[r], want[r]).real for r in range(k)])
        nerr = float(np.max(np.abs(nrm.cpu().numpy() - wn) / wn))
        print(f"in={in_prec} out={out_prec} parity={parity}: rel_err={err:.2e}, norm rel_err={nerr:.2e}")
        assert err < t
        assert nerr < 4 * t
