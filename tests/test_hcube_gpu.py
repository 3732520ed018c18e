"""GPU parity of the hypercube-mapped dslash (dslash_hcube.h) against the
FP64 oracle: every hypercube tile shape (RHS per lane, workgroup size,
block-lattice blocking, both traversals and XCD orders), N = 2..4, both
parities and dagger signs, split-lattice boundary skips, the xpay epilogue;
FP32 within 1e-6 and FP64 within 1e-13 (per-RHS relative error, max over RHS).
"""
import ctypes
import itertools

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err
from test_dslash_gpu import pack_inputs, unpack, stream, TOL

pytestmark = pytest.mark.gpu

HC = 2  # traversal bit of the hypercube mapping
COL = 4  # traversal bit of the column order (with HC: traversal 6)


def hc_tiles(L, k, blocks=(64, 128, 192, 256)):
    bt, bz = L[3] // 2, L[2] // 2
    for R in (1, 2, 4):
        if k % R or k // R > 8 or (k // R) & (k // R - 1):
            continue
        for block in blocks:
            for tb in [d for d in (1, 2, 3, 4) if bt % d == 0]:
                for zb in [d for d in (0, 1, 2, 3) if d == 0 or bz % d == 0]:
                    for ys, hints in ((0, 0), (1, 0), (1, 8)):
                        yield (R, block, 1, tb, zb, hints, 0, HC | ys)


def test_hcube_every_tile_fp32(lib, dev):
    """Every hypercube tile for N = 3, k = 8, 4, 2, 1 against the oracle."""
    L, N = (8, 4, 6, 8), 3
    g = fields.gauge_field(L, N, seed=21)
    for k in (8, 4, 2, 1):
        psi = fields.gaussian_spinors(L, N, k, seed=22 + k)
        want = {pd: oracle.dslash(g, psi, L, *pd) for pd in ((0, 0), (1, 1))}
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        n = 0
        for i, t in enumerate(hc_tiles(L, k)):
            parity, dagger = ((0, 0), (1, 1))[i & 1]
            pout = torch.zeros_like(pin)
            rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, parity, dagger,
                                   geom, _lib.tile(*t), stream())
            if rc != 0:
                msg = lib.qcuMrhsLastError().decode()
                assert "no kernel instantiated" in msg or "LDS" in msg, (t, msg)
                continue
            got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
            err = rel_err(got, want[(parity, dagger)])
            assert err < TOL[1], (k, t, err)
            n += 1
        assert n >= 20, (k, n)


@pytest.mark.parametrize("N,L,k", [(2, (4, 4, 4, 8), 4), (4, (4, 6, 4, 4), 2), (3, (2, 2, 2, 2), 2),
                                   (3, (12, 2, 4, 6), 8)])
def test_hcube_colours_and_small_extents(lib, dev, N, L, k):
    """N = 2..4 and extents down to 2 (neighbours that wrap onto the block)."""
    g = fields.gauge_field(L, N, seed=31 + N)
    psi = fields.gaussian_spinors(L, N, k, seed=32 + N)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    n = 0
    for t in hc_tiles(L, k, blocks=(64, 256)):
        for parity, dagger in ((0, 1), (1, 0)):
            pout = torch.zeros_like(pin)
            rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, parity, dagger,
                                   geom, _lib.tile(*t), stream())
            if rc != 0:
                msg = lib.qcuMrhsLastError().decode()
                assert "no kernel instantiated" in msg or "LDS" in msg, (t, msg)
                continue
            got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger)) < TOL[1], (t, parity, dagger)
            n += 1
    assert n > 0


def test_hcube_boundary_skip(lib, dev):
    """Split-lattice semantics (terms crossing a split dimension skipped)."""
    L, N, k = (8, 4, 4, 8), 3, 8
    g = fields.gauge_field(L, N, seed=41)
    psi = fields.gaussian_spinors(L, N, k, seed=42)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    for grid in ((2, 1, 1, 1), (1, 2, 2, 1), (1, 1, 1, 4), (2, 2, 2, 2)):
        for d in range(4):
            geom.grid[d] = grid[d]
        for parity in (0, 1):
            pout = torch.zeros_like(pin)
            _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, parity, 0,
                                         geom, _lib.tile(2, 128, 1, 2, 1, 0, 0, HC | 1), stream()))
            got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
            want = oracle.dslash(g, psi, L, parity, 0, grid=grid)
            assert rel_err(got, want) < TOL[1], (grid, parity)


def test_hcube_fp64(lib, dev):
    """FP64 (one RHS per lane) within 1e-13 of the oracle."""
    L, N = (4, 4, 6, 8), 3
    g = fields.gauge_field(L, N, seed=51)
    for k in (2, 4, 8):
        psi = fields.gaussian_spinors(L, N, k, seed=52 + k)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.DOUBLE)
        for parity, dagger in ((0, 0), (1, 1)):
            pout = torch.zeros_like(pin)
            _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.DOUBLE, k, parity,
                                         dagger, geom, _lib.tile(1, 64, 1, 2, 3, 0, 0, HC | 1), stream()))
            got = unpack(lib, dev, pout, _lib.DOUBLE, k, geom, psi.shape[1:])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger)) < TOL[2], (k, parity, dagger)


def test_hcube_xpay_and_k16(lib, dev):
    """out = a x + b D in through the hypercube kernel compiled for 16 RHS."""
    L, N, k = (8, 8, 4, 8), 3, 16
    g = fields.gauge_field(L, N, seed=61)
    psi = fields.gaussian_spinors(L, N, k, seed=62)
    xs = fields.gaussian_spinors(L, N, k, seed=63)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    _, _, px = pack_inputs(lib, dev, g, xs, L, N, _lib.SINGLE)
    a, b = 0.75, -0.3
    for parity, dagger in ((0, 0), (1, 1)):
        pout = torch.zeros_like(pin)
        _lib.check(lib.qcuMrhsDslashXpay(pout.data_ptr(), pin.data_ptr(), px.data_ptr(), a, b, gnat.data_ptr(),
                                         _lib.SINGLE, k, parity, dagger, geom, _lib.tile(2, 128, 1, 2, 2, 8, 0, HC | 1),
                                         stream()))
        got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
        want = a * xs + b * oracle.dslash(g, psi, L, parity, dagger)
        assert rel_err(got, want) < TOL[1], (parity, dagger)


def test_hcube_rejects_bad_tiles(lib, dev):
    """Shapes without a hypercube kernel are refused with a message."""
    L, N, k = (8, 4, 6, 8), 3, 16
    g = fields.gauge_field(L, N, seed=71)
    psi = fields.gaussian_spinors(L, N, k, seed=72)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    pout = torch.zeros_like(pin)
    for t, msg in (((1, 64, 1, 1, 0, 0, 0, HC), "multiple of 8 * nrhs"),     # 16 lanes per site
                   ((2, 128, 1, 3, 0, 0, 0, HC), "t_block must divide Lt / 2"),
                   ((2, 128, 0, 3, 0, 0, 0, HC), "t_block must divide Lt / 2"),
                   ((2, 128, 1, 1, 0, 0, 0, HC | 8), "traversal bits 3 and above"),
                   ((2, 128, 1, 1, 0, 0, 0, HC | 16 | COL), "traversal bits 3 and above"),
                   ((2, 128, 1, 1, 1, 0, 0, HC | COL | 1), "traversal 4 or 6"),
                   ((2, 128, 1, 4, 1, 0, 0, HC | COL), "column order: t_block"),
                   ((2, 128, 1, 1, 2, 0, 0, HC | COL), "column order: z_block")):
        rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, 0, 0, geom,
                               _lib.tile(*t), stream())
        assert rc != 0 and msg in lib.qcuMrhsLastError().decode(), (t, lib.qcuMrhsLastError())


def test_hcube_tile_without_kernel_runs_site_major(lib, dev):
    """A hypercube tile for a shape without a hypercube kernel (links not
    staged, N = 5, a 16-bit mixed output with run-time k) runs the site-major
    kernel of the same blocking and gives the oracle's result."""
    for N, L, k, t in ((3, (8, 4, 6, 8), 8, (2, 128, 0, 2, 3, 0, 0, HC | 1)),
                       (5, (4, 4, 4, 8), 4, (1, 128, 1, 2, 1, 8, 0, HC | 1))):
        g = fields.gauge_field(L, N, seed=81 + N)
        psi = fields.gaussian_spinors(L, N, k, seed=82 + N)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        pout = torch.zeros_like(pin)
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, 1, 0, geom,
                                     _lib.tile(*t), stream()))
        got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
        assert rel_err(got, oracle.dslash(g, psi, L, 1, 0)) < TOL[1], (N, t)


@pytest.mark.parametrize("prec", [_lib.HALF, _lib.BF16])
def test_hcube_16bit(lib, dev, prec):
    """FP16 / BF16 storage through the hypercube kernels (k = 16 compiled and
    run-time k), within the stated storage tolerance of the FP64 oracle."""
    L, N = (8, 4, 6, 8), 3
    g = fields.gauge_field(L, N, seed=91, smear_eps=0.3)
    for k, Rs in ((16, (1, 2, 4)), (8, (1, 2))):
        psi = fields.gaussian_spinors(L, N, k, seed=92 + k)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
        n = 0
        for R in Rs:
            for block in (128, 256):
                if block % (8 * (k // R)):
                    continue
                for parity, dagger in ((0, 0), (1, 1)):
                    t = (R, block, 1, 2, 3, 8 if parity else 0, 0, HC | 1)
                    pout = torch.zeros_like(pin)
                    rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger,
                                           geom, _lib.tile(*t), stream())
                    if rc != 0:
                        msg = lib.qcuMrhsLastError().decode()
                        assert "LDS" in msg, (t, msg)
                        continue
                    got = unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])
                    err = rel_err(got, oracle.dslash(g, psi, L, parity, dagger))
                    print(f"prec={prec} k={k} tile={t}: {err:.2e}")
                    assert err < TOL[prec], (k, t, err)
                    n += 1
        assert n >= 4, (k, n)


@pytest.mark.parametrize("prec", [_lib.SINGLE, _lib.DOUBLE])
def test_hcube_rhs_split_over_two_wavefronts(lib, dev, prec):
    """One RHS per lane at k = 16 (16 lanes per site: each block's slots split
    over two wavefronts), FP32 and FP64, both workgroup sizes."""
    L, N, k = (8, 4, 6, 8), 3, 16
    g = fields.gauge_field(L, N, seed=95)
    psi = fields.gaussian_spinors(L, N, k, seed=96)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
    for block in (128, 256):
        for parity, dagger, ys, hints in ((0, 0, 1, 8), (1, 1, 0, 0)):
            pout = torch.zeros_like(pin)
            rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger, geom,
                                   _lib.tile(1, block, 1, 2, 1, hints, 0, HC | ys), stream())
            if rc != 0:
                assert "LDS" in lib.qcuMrhsLastError().decode()
                continue
            got = unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger)) < TOL[prec], (block, parity)


def test_hcube_cg(lib, dev):
    """The mixed FP16 / FP64 CG with hypercube tiles on every operator it
    launches (16-bit in, FP32 out, scales, norms, the fused r update, FP32 and
    FP64 reliable updates): true residual recomputed by the oracle <= 2e-10,
    mixed solution within 1e-8 of the FP64 one."""
    from test_cg_gpu import solve, check_solution

    L, N, k, kappa = (8, 8, 4, 8), 3, 16, 0.13
    geom = _lib.geom(L, N)
    tiles = {_lib.HALF: (2, 128, 1, 2, 1, 0, 0, HC | 1), _lib.SINGLE: (2, 128, 1, 1, 2, 0, 0, HC | 1),
             _lib.DOUBLE: (1, 128, 1, 2, 1, 0, 0, HC | 1)}
    old = {}
    for prec, t in tiles.items():
        o = _lib.Tile()
        _lib.check(lib.qcuMrhsGetTile(geom, prec, k, ctypes.byref(o)))
        old[prec] = o
        _lib.check(lib.qcuMrhsSetTile(geom, prec, k, ctypes.byref(_lib.tile(*t))))
    try:
        g = fields.gauge_field(L, N, seed=97)
        b = fields.cg_sources(L, N, k, seed=98, n_point=6)
        x64, p64, r64 = solve(lib, dev, g, b, L, N, kappa, _lib.DOUBLE)
        x16, p16, r16 = solve(lib, dev, g, b, L, N, kappa, _lib.HALF)
        print(f"hypercube tiles: fp64 {p64.iterations} it, mixed {p16.iterations} it / {p16.reliable_updates} updates")
        assert max(r64) <= 1e-10 and max(r16) <= 1e-10
        assert check_solution(g, b, x16, L, kappa, 1e-10) <= 2e-10
        assert check_solution(g, b, x64, L, kappa, 1e-10) <= 2e-10
        for i in range(k):
            d = np.linalg.norm(x16[i] - x64[i]) / np.linalg.norm(x64[i])
            assert d <= 1e-8, (i, d)
    finally:
        for prec, o in old.items():
            lib.qcuMrhsSetTile(geom, prec, k, ctypes.byref(o))
        lib.qcuMrhsSolverRelease()


def test_hcube_user_layout(lib, dev):
    """The user AoS layout (qcuMrhsDslashUser, the qcu.h path) through the
    hypercube kernel, set as the user-layout tile: k = 16 and 8 (FP64 / FP32
    / FP16-with-FP32-arithmetic are not all instantiated: FP64 and FP32 here),
    both parities and dagger signs, 16-byte and unaligned (element-wise)
    buffers; k = 12 keeps the site-major pointer kernel."""
    from helpers import dev_tensor, host_complex

    L, N = (8, 4, 6, 8), 3
    g = fields.gauge_field(L, N, seed=101)
    for prec in (_lib.DOUBLE, _lib.SINGLE):
        for k, t in ((16, (1, 128, 1, 2, 3, 8, 0, HC | 1)), (8, (1, 64, 1, 1, 1, 0, 0, HC)),
                     (12, (1, 128, 1, 2, 3, 8, 0, HC | 1))):
            psi = fields.gaussian_spinors(L, N, k, seed=102 + k)
            geom, gnat, _ = pack_inputs(lib, dev, g, psi[:1], L, N, prec)
            _lib.check(lib.qcuMrhsSetTile(geom, prec | _lib.TILE_USER_LAYOUT, k, ctypes.byref(_lib.tile(*t))))
            back = _lib.Tile()
            _lib.check(lib.qcuMrhsGetTile(geom, prec | _lib.TILE_USER_LAYOUT, k, ctypes.byref(back)))
            assert back.as_tuple() == t
            cdt = np.complex128 if prec == _lib.DOUBLE else np.complex64
            for offset in (0, 1):
                # offset 1: every buffer starts one complex number into its
                # allocation (8- or 16-byte aligned only: element-wise accesses)
                n = psi[0].size * 2
                ins, outs = [], []
                for r in range(k):
                    buf = torch.zeros(n + 2, dtype=torch.float64 if prec == _lib.DOUBLE else torch.float32,
                                      device=dev)
                    buf[2 * offset:2 * offset + n] = dev_tensor(psi[r].astype(cdt), dev).reshape(-1)
                    ins.append(buf[2 * offset:2 * offset + n])
                    outs.append(torch.zeros_like(buf)[2 * offset:2 * offset + n])
                ai, k1 = _lib.ptr_array([t_.data_ptr() for t_ in ins])
                ao, k2 = _lib.ptr_array([t_.data_ptr() for t_ in outs])
                for parity, dagger in ((0, 1), (1, 0)):
                    _lib.check(lib.qcuMrhsDslashUser(ao, ai, gnat.data_ptr(), prec, k, parity, dagger, geom,
                                                     stream()))
                    torch.cuda.synchronize()
                    got = np.stack([host_complex(o.reshape(psi.shape[1:] + (2,))) for o in outs])
                    err = rel_err(got, oracle.dslash(g, psi, L, parity, dagger))
                    assert err < TOL[prec], (prec, k, offset, parity, err)


def test_hcube_tile_on_compressed_links(lib, dev):
    """Two-row and eight-real SU(3) links with a hypercube tile (the pinned
    24^3x48 FP32 tile is one, and the compressed-link entry points use the
    native layout's tile): they run the site-major kernel with that blocking
    and match the oracle on the full links."""
    from test_gauge12_gpu import run12
    from test_gauge8_gpu import run8

    L, k = (8, 4, 6, 8), 16
    g = fields.gauge_field(L, 3, seed=111)
    psi = fields.gaussian_spinors(L, 3, k, seed=112)
    t = _lib.tile(2, 128, 1, 1, 3, 8, 0, HC | 1)
    for parity, dagger in ((0, 0), (1, 1)):
        want = oracle.dslash(g, psi, L, parity, dagger)
        for run in (run12, run8):
            got = run(lib, dev, g, psi, L, parity, dagger, _lib.SINGLE, -1, tile=t)
            assert rel_err(got, want) < TOL[1], (run.__name__, parity, dagger)


def test_hcube_fp8(lib, dev):
    """FP8 (e4m3) spinors with binary16 links through the hypercube kernels
    (two and four RHS per lane): error against the FP64 oracle within the
    stated FP8 bound (0.1), and within a few FP8 roundings of the site-major
    FP8 kernel on the same input."""
    from helpers import dev_tensor, host_complex

    L, N, k = (8, 8, 8, 16), 3, 16
    geom = _lib.geom(L, N)
    g = fields.gauge_field(L, N, seed=121, smear_eps=0.3)
    psi = fields.gaussian_spinors(L, N, k, seed=122)
    gn = torch.empty(lib.qcuMrhsGaugeBytes(geom, _lib.FP8) // 2, dtype=torch.float16, device=dev)
    _lib.check(lib.qcuMrhsPackGauge(gn.data_ptr(), _lib.FP8, dev_tensor(g, dev).data_ptr(), _lib.DOUBLE, geom,
                                    stream()))
    ins = [dev_tensor(psi[r], dev) for r in range(k)]
    arr, keep = _lib.ptr_array([t.data_ptr() for t in ins])
    pin = torch.empty(lib.qcuMrhsSpinorBytes(geom, k, _lib.FP8), dtype=torch.uint8, device=dev)
    _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), _lib.FP8, arr, _lib.DOUBLE, k, geom, stream()))

    def run(t, parity, dagger):
        pout = torch.empty_like(pin)
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gn.data_ptr(), _lib.FP8, k, parity, dagger,
                                     geom, _lib.tile(*t), stream()))
        outs = [torch.empty(psi.shape[1:] + (2,), dtype=torch.float64, device=dev) for _ in range(k)]
        oarr, okeep = _lib.ptr_array([o.data_ptr() for o in outs])
        _lib.check(lib.qcuMrhsUnpackSpinors(oarr, _lib.DOUBLE, pout.data_ptr(), _lib.FP8, k, geom, stream()))
        torch.cuda.synchronize()
        return np.stack([host_complex(o) for o in outs])

    for parity, dagger in ((0, 0), (1, 1)):
        want = oracle.dslash(g, psi, L, parity, dagger)
        ref = run((2, 128, 1, 2, 0, 0, 0, 0), parity, dagger)
        for t in ((2, 128, 1, 1, 2, 8, 0, HC | 1), (4, 128, 1, 2, 1, 0, 0, HC | 1)):
            got = run(t, parity, dagger)
            err = rel_err(got, want)
            print(f"FP8 hypercube tile {t}, parity {parity}: error vs FP64 oracle {err:.3e}, "
                  f"vs site-major FP8 {rel_err(got, ref):.3e}")
            assert err < 0.1 and rel_err(got, ref) < 0.03, (t, parity)


def col_tiles(L, k, Rs=(1, 2, 4), blocks=(64, 128, 256)):
    """Column-order tiles (traversal 6) whose column slabs fill whole workgroups."""
    by, bz = L[1] // 2, L[2] // 2
    for R in Rs:
        P = k // R if k % R == 0 else 0
        if not P or P > 32 or P & (P - 1):
            continue
        for block in blocks:
            if block % (8 * P):
                continue
            for cy in [d for d in (1, 2, 3, 4) if by % d == 0]:
                for cz in [d for d in (1, 2, 3, 6) if bz % d == 0]:
                    if (L[0] // 2 * cy * cz) % (block // (8 * P)) == 0:
                        yield (R, block, 1, cy, cz, 0, 0, HC | COL)
                    if (L[0] // 4 * cy * cz) % (block // (8 * P)) == 0 and L[0] % 4 == 0:
                        yield (R, block, 1, cy, cz, 0, 0, HC | COL | 8)  # x split in two


def test_column_order_every_tile_fp32(lib, dev):
    """Column order: XCD x walks columns 8 g + x along t, the last group of 8
    padded with idle workgroups (1 to 24 columns here). Every column shape,
    RHS per lane and workgroup size for k = 16, 8, 2 against the oracle."""
    L, N = (8, 8, 12, 8), 3
    g = fields.gauge_field(L, N, seed=131)
    for k in (16, 8, 2):
        psi = fields.gaussian_spinors(L, N, k, seed=132 + k)
        want = {pd: oracle.dslash(g, psi, L, *pd) for pd in ((0, 0), (1, 1))}
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        n = 0
        for i, t in enumerate(col_tiles(L, k)):
            parity, dagger = ((0, 0), (1, 1))[i & 1]
            pout = torch.full_like(pin, float("nan"))
            rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), _lib.SINGLE, k, parity, dagger,
                                   geom, _lib.tile(*t), stream())
            if rc != 0:
                msg = lib.qcuMrhsLastError().decode()
                assert "LDS" in msg, (t, msg)
                continue
            got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
            err = rel_err(got, want[(parity, dagger)])
            assert err < TOL[1], (k, t, err)
            n += 1
        assert n >= 10, (k, n)


@pytest.mark.parametrize("prec,k,R", [(_lib.DOUBLE, 4, 1), (_lib.HALF, 16, 4), (_lib.BF16, 16, 2),
                                      (_lib.HALF, 8, 2)])
def test_column_order_precisions(lib, dev, prec, k, R):
    """Column order with FP64, FP16 and BF16 storage, boundary skips on a
    split lattice included."""
    L, N = (8, 4, 12, 8), 3
    g = fields.gauge_field(L, N, seed=141)
    psi = fields.gaussian_spinors(L, N, k, seed=142 + k)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
    tol = {_lib.HALF: 3e-3, _lib.BF16: 2e-2, _lib.DOUBLE: TOL[2]}[prec]
    n = 0
    for grid in ((1, 1, 1, 1), (1, 1, 2, 2)):
        for d in range(4):
            geom.grid[d] = grid[d]
        for i, t in enumerate(col_tiles(L, k, Rs=(R,), blocks=(128,))):
            parity, dagger = ((0, 1), (1, 0))[i & 1]
            pout = torch.zeros_like(pin)
            _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger,
                                         geom, _lib.tile(*t), stream()))
            got = unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])
            assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger, grid=grid)) < tol, (grid, t)
            n += 1
    assert n >= 8


def test_column_order_norms_and_fallbacks(lib, dev):
    """A column-order tile set for a shape serves qcuMrhsDslashEx (per-RHS
    norms from per-workgroup partials indexed by the logical workgroup); on
    calls no hypercube kernel serves (two-row and eight-real links, N = 5) it
    runs the default traversal; both against the oracle."""
    from test_gauge12_gpu import run12
    from test_gauge8_gpu import run8

    L, N, k = (8, 8, 12, 8), 3, 16
    g = fields.gauge_field(L, N, seed=151)
    psi = fields.gaussian_spinors(L, N, k, seed=152)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    t = _lib.tile(2, 128, 1, 2, 3, 0, 0, HC | COL)
    _lib.check(lib.qcuMrhsSetTile(geom, _lib.SINGLE, k, ctypes.byref(t)))
    pout = torch.zeros_like(pin)
    nrm = torch.zeros(k, dtype=torch.float64, device=dev)
    b = -0.25
    for parity, dagger in ((0, 1), (1, 0)):
        _lib.check(lib.qcuMrhsDslashEx(pout.data_ptr(), _lib.SINGLE, pin.data_ptr(), None, _lib.SINGLE, 0.0, b,
                                       None, None, None, nrm.data_ptr(), gnat.data_ptr(), k, parity, dagger, geom,
                                       stream()))
        got = unpack(lib, dev, pout, _lib.SINGLE, k, geom, psi.shape[1:])
        want = b * oracle.dslash(g, psi, L, parity, dagger)
        assert rel_err(got, want) < TOL[1], (parity, dagger)
        wn = np.array([np.vdot(want[r], want[r]).real for r in range(k)])
        assert np.max(np.abs(nrm.cpu().numpy() - wn) / wn) < 4e-6, (parity, dagger)
    for run in (run12, run8):
        got = run(lib, dev, g, psi, L, 1, 0, _lib.SINGLE, -1, tile=t)
        assert rel_err(got, oracle.dslash(g, psi, L, 1, 0)) < TOL[1], run.__name__
    L5, N5, k5 = (4, 4, 4, 8), 5, 4
    g5 = fields.gauge_field(L5, N5, seed=153)
    psi5 = fields.gaussian_spinors(L5, N5, k5, seed=154)
    geom5, gnat5, pin5 = pack_inputs(lib, dev, g5, psi5, L5, N5, _lib.SINGLE)
    pout5 = torch.zeros_like(pin5)
    _lib.check(lib.qcuMrhsDslash(pout5.data_ptr(), pin5.data_ptr(), gnat5.data_ptr(), _lib.SINGLE, k5, 0, 0, geom5,
                                 _lib.tile(1, 128, 1, 1, 1, 0, 0, HC | COL), stream()))
    got = unpack(lib, dev, pout5, _lib.SINGLE, k5, geom5, psi5.shape[1:])
    assert rel_err(got, oracle.dslash(g5, psi5, L5, 0, 0)) < TOL[1]


@pytest.mark.parametrize("N,L,k,R,prec", [(3, (8, 8, 12, 8), 16, 2, _lib.SINGLE), (3, (8, 8, 12, 8), 12, 2, _lib.SINGLE),
                                          (2, (8, 4, 8, 8), 8, 1, _lib.SINGLE), (4, (4, 8, 4, 8), 6, 2, _lib.SINGLE),
                                          (5, (4, 4, 8, 8), 4, 1, _lib.SINGLE), (3, (8, 8, 12, 8), 4, 1, _lib.DOUBLE),
                                          (3, (8, 8, 12, 8), 16, 4, _lib.HALF)])
def test_site_major_column_order(lib, dev, N, L, k, R, prec):
    """Column order in sites (traversal 4, the site-major kernels): every
    column shape whose columns fill whole workgroups, N = 2..5, k = 4..16,
    FP32 / FP64 / FP16, a split lattice included, against the oracle;
    shapes whose columns do not fill whole workgroups run y-streamed blocks."""
    g = fields.gauge_field(L, N, seed=161 + N)
    psi = fields.gaussian_spinors(L, N, k, seed=162 + k)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, prec)
    tol = {_lib.HALF: 3e-3, _lib.SINGLE: TOL[1], _lib.DOUBLE: TOL[2]}[prec]
    n = 0
    for grid in ((1, 1, 1, 1), (1, 2, 1, 2)):
        for d in range(4):
            geom.grid[d] = grid[d]
        i = 0
        for cy in [d for d in (1, 2, 4, 8) if L[1] % d == 0]:
            for cz in [d for d in (1, 2, 3, 4, 12) if L[2] % d == 0]:
                for block, lds in ((128, 1), (256, 0)):
                    parity, dagger = ((0, 1), (1, 0))[i & 1]
                    i += 1
                    t = (R, block, lds, cy, cz, 0, 0, COL)
                    pout = torch.full_like(pin, float("nan"))
                    rc = lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger,
                                           geom, _lib.tile(*t), stream())
                    if rc != 0:
                        msg = lib.qcuMrhsLastError().decode()
                        assert "LDS" in msg or "no kernel instantiated" in msg, (t, msg)
                        continue
                    got = unpack(lib, dev, pout, prec, k, geom, psi.shape[1:])
                    assert rel_err(got, oracle.dslash(g, psi, L, parity, dagger, grid=grid)) < tol, (grid, t)
                    n += 1
    assert n >= 8
