"""CPU oracle: Python access to the C restatement of the reference dslash
(oracle.c) plus numpy restatements of the reference's layout shifts.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check of
__graft_entry__.py and by the cpu_baseline leg of bench.py, never by the
library.

Functions and the reference lines they restate:
  dslash            include/kernels/wilson_dslash_4d.cuh:27-301 (via oracle.c)
  shift_vector      include/kernels/storage/storage_kernels.cuh:14-52
  shift_gauge       include/kernels/storage/storage_kernels.cuh:54-101
  ref_kernel_*      the reference's own kernels compiled into _ref/ (GPU)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None


def load():
    global _lib
    if _lib is None:
        p = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        lib = ctypes.CDLL(p)
        for name in ("oracle_dslash_f64", "oracle_dslash_f32"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
        _lib = lib
    return _lib


def _ints(v):
    return (ctypes.c_int * 4)(*[int(a) for a in v])


def dslash(gauge, psi, L, parity, dagger, grid=(1, 1, 1, 1), single=False, out=None):
    """Hopping term on k user-layout spinors.

    gauge: [4][2][S][N][N] complex; psi: [k][S][4][N] complex (parity 1-parity);
    returns [k][S][4][N] on `parity`. single=True runs the FP32 restatement.
    """
    lib = load()
    k, S, four, N = psi.shape
    assert four == 4 and gauge.shape == (4, 2, S, N, N)
    assert k <= 64 and N <= 8
    dt = np.complex64 if single else np.complex128
    g = np.ascontiguousarray(gauge, dtype=dt)
    x = np.ascontiguousarray(psi, dtype=dt)
    if out is None:
        out = np.empty_like(x)
    fn = lib.oracle_dslash_f32 if single else lib.oracle_dslash_f64
    fn(out.ctypes.data, x.ctypes.data, g.ctypes.data, k, N, _ints(L), _ints(grid), int(parity), int(dagger))
    return out


def dslash_many(gauge, psi, L, parity, dagger, grid=(1, 1, 1, 1), single=False):
    """dslash for any number of RHS (chunks of 64)."""
    out = np.empty_like(psi, dtype=np.complex64 if single else np.complex128)
    for a in range(0, psi.shape[0], 64):
        out[a:a + 64] = dslash(gauge, psi[a:a + 64], L, parity, dagger, grid, single)
    return out


def shift_vector(v, to_soa, S, N):
    """storeVectorCoalesced / storeVectorNonCoalesced (storage_kernels.cuh:14-52):
    dst[e*S + s] = src[s*4N + e] and its inverse, as pure index maps."""
    M = 4 * N
    v = np.asarray(v).reshape(-1)
    out = np.empty_like(v)
    s = np.arange(S)[:, None]
    e = np.arange(M)[None, :]
    if to_soa:
        out[(e * S + s).ravel()] = v[(s * M + e).ravel()]
    else:
        out[(s * M + e).ravel()] = v[(e * S + s).ravel()]
    return out


def shift_gauge(g, to_soa, S, N):
    """storeGaugeCoalesced / storeGaugeNonCoalesced (storage_kernels.cuh:54-101):
    for each (dir, parity) block b: dst[b*S*N^2 + j*S + s] = src[b*S*N^2 + s*N^2 + j]."""
    M = N * N
    g = np.asarray(g).reshape(-1)
    out = np.empty_like(g)
    s = np.arange(S)[:, None]
    j = np.arange(M)[None, :]
    for b in range(8):
        base = b * S * M
        if to_soa:
            out[base + (j * S + s).ravel()] = g[base + (s * M + j).ravel()]
        else:
            out[base + (s * M + j).ravel()] = g[base + (j * S + s).ravel()]
    return out


def ref_kernels():
    """The reference's own kernels (built from its sources into _ref/); None
    when that build is absent."""
    global _ref
    if _ref is None:
        p = os.path.join(_HERE, "_ref", "libqcu_ref.so")
        if not os.path.exists(p):
            return None
        lib = ctypes.CDLL(p)
        PP = ctypes.POINTER(ctypes.c_void_p)
        lib.ref_dslash.restype = ctypes.c_int
        lib.ref_dslash.argtypes = [PP, PP, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                   ctypes.c_int]
        for n in ("ref_shift_vector", "ref_shift_gauge"):
            f = getattr(lib, n)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 5
        _ref = lib
    return _ref
