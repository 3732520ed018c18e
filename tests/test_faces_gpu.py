"""Domain decomposition on one GPU: a global lattice is split into
sub-lattices (as processes would hold them), each applies the interior
operator (qcuMrhsDslash with grid > 1: boundary terms skipped), packs its
faces (qcuMrhsPackFaces), takes its neighbours' faces (here: their device
buffers, which a halo exchange would move) and adds the boundary terms
(qcuMrhsDslashFaces). The gathered result must equal the FP64 oracle's
operator on the whole periodic lattice: FP64 <= 1e-13, FP32 <= 1e-6,
FP16 storage <= 3e-3 (measured error printed)."""
import ctypes
import itertools

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err, dev_tensor, host_complex, empty_prec, cb_index

pytestmark = pytest.mark.gpu

TOL = {2: 1e-13, 1: 1e-6, 0: 3e-3}


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def split_run(lib, dev, L, P, N, k, prec, parity, dagger, seed=0):
    g = fields.gauge_field(L, N, seed=seed)
    psi = fields.gaussian_spinors(L, N, k, seed=seed + 1)
    want = oracle.dslash(g, psi, L, parity, dagger)
    Ll = [L[d] // P[d] for d in range(4)]
    geom = _lib.geom(Ll, N, grid=P)
    xh, y, z, t = fields.site_coords(Ll)
    ranks = list(itertools.product(*[range(p) for p in P]))

    def global_index(rk, par):
        X = 2 * xh + ((par + y + z + t) & 1)
        return cb_index(L, X + rk[0] * Ll[0], y + rk[1] * Ll[1], z + rk[2] * Ll[2], t + rk[3] * Ll[3])

    q = 1 - parity
    st = stream()
    keep = []
    state = {}
    for rk in ranks:
        gl = np.stack([g[:, par, global_index(rk, par)] for par in (0, 1)], axis=1)
        gd = dev_tensor(gl, dev)
        gnat = empty_prec((lib.qcuMrhsGaugeBytes(geom, prec) // _lib.BYTES[prec],), prec, dev)
        _lib.check(lib.qcuMrhsPackGauge(gnat.data_ptr(), prec, gd.data_ptr(), _lib.DOUBLE, geom, st))
        ins = [dev_tensor(psi[r][global_index(rk, q)], dev) for r in range(k)]
        arr, a_keep = _lib.ptr_array([x.data_ptr() for x in ins])
        pin = empty_prec((fields.half_volume(Ll), 4, N, k), prec, dev)
        _lib.check(lib.qcuMrhsPackSpinors(pin.data_ptr(), prec, arr, _lib.DOUBLE, k, geom, st))
        pout = torch.empty_like(pin)
        _lib.check(lib.qcuMrhsDslash(pout.data_ptr(), pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger, geom,
                                     None, st))
        to_bwd, to_fwd = [0] * 4, [0] * 4
        for mu in range(4):
            if P[mu] > 1:
                nb = lib.qcuMrhsFaceBytes(geom, k, prec, mu)
                assert nb > 0
                tb = empty_prec((nb // _lib.BYTES[prec],), prec, dev)
                tf = empty_prec((nb // _lib.BYTES[prec],), prec, dev)
                keep += [tb, tf]
                to_bwd[mu], to_fwd[mu] = tb.data_ptr(), tf.data_ptr()
        pb, kb = _lib.ptr_array(to_bwd)
        pf, kf = _lib.ptr_array(to_fwd)
        _lib.check(lib.qcuMrhsPackFaces(pb, pf, pin.data_ptr(), gnat.data_ptr(), prec, k, parity, dagger, geom, st))
        keep += [gd, ins, a_keep, kb, kf]
        state[rk] = (gnat, pin, pout, to_bwd, to_fwd)
    got = np.zeros_like(want)
    for rk in ranks:
        gnat, pin, pout, _, _ = state[rk]
        from_bwd, from_fwd = [0] * 4, [0] * 4
        for mu in range(4):
            if P[mu] > 1:
                ahead = list(rk)
                ahead[mu] = (rk[mu] + 1) % P[mu]
                behind = list(rk)
                behind[mu] = (rk[mu] - 1) % P[mu]
                from_fwd[mu] = state[tuple(ahead)][3][mu]
                from_bwd[mu] = state[tuple(behind)][4][mu]
        pb, kb = _lib.ptr_array(from_bwd)
        pf, kf = _lib.ptr_array(from_fwd)
        _lib.check(lib.qcuMrhsDslashFaces(pout.data_ptr(), pb, pf, gnat.data_ptr(), prec, k, parity, dagger, geom,
                                          st))
        outs = [torch.empty((fields.half_volume(Ll), 4, N, 2), dtype=torch.float64, device=dev) for _ in range(k)]
        arr, a_keep = _lib.ptr_array([x.data_ptr() for x in outs])
        _lib.check(lib.qcuMrhsUnpackSpinors(arr, _lib.DOUBLE, pout.data_ptr(), prec, k, geom, st))
        torch.cuda.synchronize()
        got[:, global_index(rk, parity)] = np.stack([host_complex(x) for x in outs])
    return rel_err(got, want)


@pytest.mark.parametrize("P", [(2, 1, 1, 1), (1, 2, 1, 1), (1, 1, 2, 1), (1, 1, 1, 2), (1, 1, 2, 2), (2, 2, 2, 2)])
@pytest.mark.parametrize("prec", [_lib.DOUBLE, _lib.SINGLE])
def test_split_lattice_matches_whole(lib, dev, P, prec):
    """SU(3), 4 RHS, global 8x4x4x8 (8^4 for the 16-way split), both parities
    and dagger signs."""
    L = (8, 8, 8, 8) if P == (2, 2, 2, 2) else (8, 4, 4, 8)
    for parity, dagger in ((0, 0), (1, 0), (0, 1), (1, 1)):
        err = split_run(lib, dev, L, P, 3, 4, prec, parity, dagger)
        print(f"split {P} prec {prec} parity {parity} dagger {dagger}: rel_err {err:.3e}")
        assert err < TOL[prec]


@pytest.mark.parametrize("N", [2, 4, 5])
def test_split_other_colours(lib, dev, N):
    """SU(2), SU(4), SU(5) with t and x split, FP32, 3 RHS."""
    err = split_run(lib, dev, (8, 4, 4, 8), (2, 1, 1, 2), N, 3, _lib.SINGLE, 1, 0, seed=5)
    print(f"N={N}: rel_err {err:.3e}")
    assert err < TOL[_lib.SINGLE]


def test_split_fp16_storage(lib, dev):
    """FP16 spinors, links and faces with FP32 arithmetic: stated error."""
    err = split_run(lib, dev, (8, 4, 4, 8), (1, 1, 2, 2), 3, 8, _lib.HALF, 0, 0)
    print(f"fp16 split: rel_err {err:.3e}")
    assert err < TOL[_lib.HALF]


def test_no_split_dimension_is_a_no_op(lib, dev):
    """With every grid entry 1 nothing is packed or added."""
    L, N, k = (4, 4, 4, 4), 3, 2
    geom = _lib.geom(L, N)
    S = fields.half_volume(L)
    out = torch.randn((S * 4 * N * k * 2,), dtype=torch.float32, device=dev)
    before = out.clone()
    gauge = torch.zeros((S * 2 * 8 * N * N * 2,), dtype=torch.float32, device=dev)
    null4, keep = _lib.ptr_array([0, 0, 0, 0])
    _lib.check(lib.qcuMrhsPackFaces(null4, null4, out.data_ptr(), gauge.data_ptr(), _lib.SINGLE, k, 0, 0, geom,
                                    stream()))
    _lib.check(lib.qcuMrhsDslashFaces(out.data_ptr(), null4, null4, gauge.data_ptr(), _lib.SINGLE, k, 0, 0, geom,
                                      stream()))
    torch.cuda.synchronize()
    assert torch.equal(out, before)
