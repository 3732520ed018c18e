"""ctypes binding of libqcu.so (include/qcu.h + include/qcu_mrhs.h).

The library is the product; this module only declares argument types. It
raises ``RuntimeError`` when the library is missing: there is no fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqcu.so")

HALF, SINGLE, DOUBLE, BF16, FP8 = 0, 1, 2, 4, 5
TILE_USER_LAYOUT = 8  # precision flag of GetTile / SetTile: the user-layout (AoS pointer) path
BYTES = {HALF: 4, SINGLE: 8, DOUBLE: 16, BF16: 4, FP8: 2}  # per complex (FP8 links are stored as HALF)


class QcuParam(ctypes.Structure):
    _fields_ = [("lattice_size", ctypes.c_int * 4)]


class QcuGrid(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int * 4)]


class Geom(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int * 4), ("grid", ctypes.c_int * 4), ("nColors", ctypes.c_int)]


class Tile(ctypes.Structure):
    _fields_ = [("rhs_per_lane", ctypes.c_int), ("block", ctypes.c_int), ("link_lds", ctypes.c_int),
                ("t_block", ctypes.c_int), ("z_block", ctypes.c_int), ("cache_hints", ctypes.c_int),
                ("occupancy", ctypes.c_int), ("traversal", ctypes.c_int)]

    def as_tuple(self):
        return (self.rhs_per_lane, self.block, self.link_lds, self.t_block, self.z_block, self.cache_hints,
                self.occupancy, self.traversal)


class CgParam(ctypes.Structure):
    _fields_ = [("kappa", ctypes.c_double), ("tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("reliable_interval", ctypes.c_int), ("sloppy_prec", ctypes.c_int),
                ("reliable_delta", ctypes.c_double), ("iterations", ctypes.c_int), ("reliable_updates", ctypes.c_int), ("seconds", ctypes.c_double)]


class Problem(ctypes.Structure):
    _fields_ = [("geom", Geom), ("nrhs", ctypes.c_int), ("parity", ctypes.c_int), ("dagger", ctypes.c_int),
                ("out", ctypes.c_void_p), ("in_", ctypes.c_void_p), ("gauge", ctypes.c_void_p)]


# Every extern "C" symbol include/qcu.h and include/qcu_mrhs.h declare.
QCU_H_SYMBOLS = ["initGridSize", "pushBackFermions", "startDslash", "loadQcuGauge", "getDslash", "dslashQcu",
                 "fullDslashQcu", "cg_inverter"]
QCU_MRHS_H_SYMBOLS = ["qcuMrhsLastError", "qcuMrhsVersion", "qcuMrhsSpinorBytes", "qcuMrhsGaugeBytes",
                      "qcuMrhsPackGauge", "qcuMrhsPackSpinors", "qcuMrhsUnpackSpinors", "qcuMrhsConvert",
                      "qcuMrhsShiftSpinor", "qcuMrhsShiftGauge", "qcuMrhsDslash", "qcuMrhsDslashXpay", "qcuMrhsDslashEx",
                      "qcuMrhsDslashUser", "qcuMrhsDslashReferenceLayout", "qcuMrhsTune", "qcuMrhsGetTile",
                      "qcuMrhsKernelCount", "qcuMrhsCgSolve", "qcuMrhsSolverRelease", "qcuMrhsBatchCreate", "qcuMrhsBatchRun", "qcuMrhsBatchGetTile",
                      "qcuMrhsBatchDestroy", "qcuMrhsFaceBytes", "qcuMrhsPackFaces", "qcuMrhsDslashFaces",
                      "qcuMrhsGauge12Bytes", "qcuMrhsPackGauge12", "qcuMrhsDslash12",
                      "qcuMrhsGauge8Bytes", "qcuMrhsPackGauge8", "qcuMrhsDslash8",
                      "qcuMrhsCloverBytes", "qcuMrhsCloverCompute", "qcuMrhsCloverApply",
                      "qcuMrhsCgSolveClover", "qcuMrhsStreamBandwidth", "qcuMrhsDslashExNoAlloc",
                      "qcuMrhsLinkBlockBench", "qcuMrhsSetTile"]

_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
PP = ctypes.POINTER(ctypes.c_void_p)


def _declare(lib):
    G = ctypes.POINTER(Geom)
    T = ctypes.POINTER(Tile)
    sig = {
        "initGridSize": (None, [ctypes.POINTER(QcuGrid), ctypes.POINTER(QcuParam), I]),
        "pushBackFermions": (None, [P, P]),
        "startDslash": (None, [I, I]),
        "loadQcuGauge": (None, [P]),
        "getDslash": (None, [I, D, I, I, I]),
        "dslashQcu": (None, [P, P, P, ctypes.POINTER(QcuParam), I]),
        "fullDslashQcu": (None, [P, P, P, ctypes.POINTER(QcuParam), I]),
        "cg_inverter": (None, [P, P, P, ctypes.POINTER(QcuParam), D, D]),
        "qcuMrhsLastError": (ctypes.c_char_p, []),
        "qcuMrhsVersion": (I, []),
        "qcuMrhsSpinorBytes": (ctypes.c_size_t, [G, I, I]),
        "qcuMrhsGaugeBytes": (ctypes.c_size_t, [G, I]),
        "qcuMrhsPackGauge": (I, [P, I, P, I, G, P]),
        "qcuMrhsPackSpinors": (I, [P, I, PP, I, I, G, P]),
        "qcuMrhsUnpackSpinors": (I, [P, I, PP, I, I, G, P]),
        "qcuMrhsConvert": (I, [P, I, P, I, ctypes.c_size_t, P]),
        "qcuMrhsShiftSpinor": (I, [P, P, I, G, I, P]),
        "qcuMrhsShiftGauge": (I, [P, P, I, G, I, P]),
        "qcuMrhsDslash": (I, [P, P, P, I, I, I, I, G, T, P]),
        "qcuMrhsDslashXpay": (I, [P, P, P, D, D, P, I, I, I, I, G, T, P]),
        "qcuMrhsDslashUser": (I, [PP, PP, P, I, I, I, I, G, P]),
        "qcuMrhsDslashEx": (I, [P, I, P, P, I, D, D, P, P, P, P, P, I, I, I, G, P]),
        "qcuMrhsDslashExNoAlloc": (I, [P, I, P, P, I, D, D, P, P, P, P, P, I, I, I, G, P]),
        "qcuMrhsDslashReferenceLayout": (I, [PP, PP, P, I, I, I, I, G]),
        "qcuMrhsTune": (I, [P, P, P, I, I, G, T, ctypes.POINTER(D), P]),
        "qcuMrhsGetTile": (I, [G, I, I, T]),
        "qcuMrhsSetTile": (I, [G, I, I, T]),
        "qcuMrhsKernelCount": (I, []),
        "qcuMrhsCgSolve": (I, [P, P, P, I, G, ctypes.POINTER(CgParam), ctypes.POINTER(D), P]),
        "qcuMrhsSolverRelease": (I, []),
        "qcuMrhsBatchCreate": (I, [ctypes.POINTER(Problem), I, ctypes.POINTER(P), P]),
        "qcuMrhsBatchRun": (I, [P, P]),
        "qcuMrhsBatchGetTile": (I, [P, I, T]),
        "qcuMrhsBatchDestroy": (I, [P]),
        "qcuMrhsFaceBytes": (ctypes.c_size_t, [G, I, I, I]),
        "qcuMrhsPackFaces": (I, [PP, PP, P, P, I, I, I, I, G, P]),
        "qcuMrhsDslashFaces": (I, [P, PP, PP, P, I, I, I, I, G, P]),
        "qcuMrhsGauge12Bytes": (ctypes.c_size_t, [G, I]),
        "qcuMrhsPackGauge12": (I, [P, I, P, I, G, I, P]),
        "qcuMrhsDslash12": (I, [P, P, P, I, I, I, I, I, G, T, P]),
        "qcuMrhsGauge8Bytes": (ctypes.c_size_t, [G, I]),
        "qcuMrhsPackGauge8": (I, [P, I, P, I, G, I, P]),
        "qcuMrhsDslash8": (I, [P, P, P, I, I, I, I, I, G, T, P]),
        "qcuMrhsCloverBytes": (ctypes.c_size_t, [G, I]),
        "qcuMrhsCloverCompute": (I, [P, P, I, P, I, D, G, P]),
        "qcuMrhsCloverApply": (I, [P, P, P, I, I, I, G, P]),
        "qcuMrhsCgSolveClover": (I, [P, P, P, P, P, I, G, ctypes.POINTER(CgParam), ctypes.POINTER(D), P]),
        "qcuMrhsStreamBandwidth": (I, [I, P, P, ctypes.c_size_t, I, ctypes.POINTER(D), P]),
        "qcuMrhsLinkBlockBench": (I, [I, I, I, I, ctypes.POINTER(D), ctypes.POINTER(D), P]),
    }
    # unpack takes a pointer array as its first argument
    sig["qcuMrhsUnpackSpinors"] = (I, [PP, I, P, I, I, G, P])
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def load(path=None):
    """Load libqcu.so once; raise RuntimeError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built: run __graft_entry__.build() (make -C mrhs_qcu_amd/csrc)")
    lib = ctypes.CDLL(p)
    _declare(lib)
    _lib = lib
    return lib


def check(rc):
    """Raise with the library's message when an extended entry point fails."""
    if rc != 0:
        raise RuntimeError(load().qcuMrhsLastError().decode())


def geom(L, N, grid=(1, 1, 1, 1)):
    g = Geom()
    for d in range(4):
        g.L[d] = int(L[d])
        g.grid[d] = int(grid[d])
    g.nColors = int(N)
    return g


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
    return ctypes.cast(arr, PP), arr


def tile(R=1, block=256, lds=0, tblock=1, zblock=0, hints=0, occ=0, traversal=0):
    t = Tile()
    t.rhs_per_lane, t.block, t.link_lds, t.t_block, t.z_block, t.cache_hints = R, block, lds, tblock, zblock, hints
    t.occupancy, t.traversal = occ, traversal
    return t
