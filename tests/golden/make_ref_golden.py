"""Writes golden vectors from the REFERENCE's own dslash and shift kernels
(oracle/_ref/libqcu_ref.so, built from the reference sources by oracle/Makefile)
on a GPU: python tests/golden/make_ref_golden.py <out_dir>.

For every case of ref_cases.CASES the inputs come from the seeded generators;
the fixture stores the input SHA-256 and the reference kernel's output in the
user layout [k][S][4][N]. The shift fixture stores the reference
storeVectorCoalesced / storeGaugeCoalesced outputs for a 4^4 SU(3) field.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import ref_cases  # noqa: E402
from oracle import oracle  # noqa: E402


def main(out_dir):
    import ctypes

    import torch

    from mrhs_qcu_amd import fields

    ref = oracle.ref_kernels()
    assert ref is not None, "oracle/_ref/libqcu_ref.so missing"
    dev = torch.device("cuda:0")
    os.makedirs(out_dir, exist_ok=True)
    out = {}
    for case in ref_cases.CASES:
        name, L, N, k, prec, parity, dagger, grid, seed = case
        print("case", name, flush=True)
        g, psi = ref_cases.inputs(case)
        S = fields.half_volume(L)
        gsoa = oracle.shift_gauge(g.reshape(-1), True, S, N)
        gd = torch.from_numpy(gsoa.view(np.float64 if prec == 2 else np.float32).copy()).to(dev)
        ins = [torch.from_numpy(oracle.shift_vector(psi[r].reshape(-1), True, S, N)
                                .view(np.float64 if prec == 2 else np.float32).copy()).to(dev) for r in range(k)]
        outs = [torch.empty_like(t) for t in ins]
        PP = ctypes.c_void_p * k
        ai = PP(*[t.data_ptr() for t in ins])
        ao = PP(*[t.data_ptr() for t in outs])
        rc = ref.ref_dslash(ctypes.cast(ao, ctypes.POINTER(ctypes.c_void_p)),
                            ctypes.cast(ai, ctypes.POINTER(ctypes.c_void_p)), gd.data_ptr(), k, N, prec,
                            (ctypes.c_int * 4)(*L), (ctypes.c_int * 4)(*grid), parity, dagger)
        assert rc == 0, rc
        res = []
        for t in outs:
            a = t.cpu().numpy()
            c = a.view(np.complex128 if prec == 2 else np.complex64)
            res.append(oracle.shift_vector(c, False, S, N).reshape(S, 4, N))
        out[name + "__out"] = np.stack(res)
        out[name + "__sha"] = np.array(ref_cases.digest(g, psi))
    # shift kernels
    L, N = (4, 4, 4, 4), 3
    S = fields.half_volume(L)
    v = fields.gaussian_spinors(L, N, 1, seed=7)[0].reshape(-1)
    g = fields.gauge_field(L, N, seed=8).reshape(-1)
    vd = torch.from_numpy(v.view(np.float64).copy()).to(dev)
    gd = torch.from_numpy(g.view(np.float64).copy()).to(dev)
    vs, gs = torch.empty_like(vd), torch.empty_like(gd)
    assert ref.ref_shift_vector(vs.data_ptr(), vd.data_ptr(), 0, L[0] // 2, L[1], L[2], L[3]) == 0
    assert ref.ref_shift_gauge(gs.data_ptr(), gd.data_ptr(), 0, L[0], L[1], L[2], L[3]) == 0
    out["shift_vector_soa"] = vs.cpu().numpy().view(np.complex128)
    out["shift_gauge_soa"] = gs.cpu().numpy().view(np.complex128)
    out["shift_sha"] = np.array(ref_cases.digest(v, g))
    path = os.path.join(out_dir, "ref_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.getcwd())
