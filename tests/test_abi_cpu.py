"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare, validates arguments without touching the GPU, and keeps the
reference's fatal-error convention (message on stderr, exit status 1)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|size_t|char)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol(lib):
    from mrhs_qcu_amd import _lib

    for header, listed in (("qcu.h", _lib.QCU_H_SYMBOLS), ("qcu_mrhs.h", _lib.QCU_MRHS_H_SYMBOLS)):
        names = declared_functions(header)
        assert sorted(names) == sorted(listed), (header, names)
        for n in names:
            assert hasattr(lib, n), n


def test_kernel_registry(lib):
    assert lib.qcuMrhsKernelCount() >= 120
    assert lib.qcuMrhsVersion() >= 10000


def test_sizes(lib):
    from mrhs_qcu_amd import _lib

    g = _lib.geom((24, 24, 24, 48), 3)
    S = 12 * 24 * 24 * 48
    assert lib.qcuMrhsSpinorBytes(g, 16, _lib.SINGLE) == S * 12 * 16 * 8
    assert lib.qcuMrhsGaugeBytes(g, _lib.HALF) == S * 2 * 8 * 9 * 4
    bad = _lib.geom((24, 24, 23, 48), 3)
    assert lib.qcuMrhsSpinorBytes(bad, 16, _lib.SINGLE) == 0


def test_argument_validation_without_gpu(lib):
    from mrhs_qcu_amd import _lib

    t = _lib.tile()
    cases = [
        (_lib.geom((4, 4, 4, 5), 3), 2, "even"),
        (_lib.geom((4, 4, 4, 4), 9), 2, "nColors"),
        (_lib.geom((4, 4, 4, 4), 3), 0, "nrhs"),
    ]
    for g, k, msg in cases:
        rc = lib.qcuMrhsDslash(16, 16, 16, _lib.SINGLE, k, 0, 0, g, t, None)
        assert rc != 0
        assert msg in lib.qcuMrhsLastError().decode()
    g = _lib.geom((4, 4, 4, 4), 3)
    assert lib.qcuMrhsDslash(16, 16, 16, _lib.SINGLE, 3, 0, 0, g, _lib.tile(R=2), None) != 0
    assert "multiple" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash(16, 16, 16, _lib.SINGLE, 2, 0, 0, g, _lib.tile(block=96), None) != 0
    assert lib.qcuMrhsDslash(8, 16, 16, _lib.SINGLE, 2, 0, 0, g, _lib.tile(), None) != 0
    assert "aligned" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash(16, 16, 16, 3, 2, 0, 0, g, _lib.tile(), None) != 0
    assert lib.qcuMrhsDslash(16, 16, 16, _lib.SINGLE, 2, 2, 0, g, _lib.tile(), None) != 0
    h = ctypes.c_void_p()
    assert lib.qcuMrhsBatchCreate(None, 0, ctypes.byref(h), None) != 0
    assert "at least one problem" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsBatchRun(None, None) != 0 and lib.qcuMrhsBatchDestroy(None) != 0
    assert "null batch handle" in lib.qcuMrhsLastError().decode()


def test_gauge12_sizes_and_validation(lib):
    """Two-row links: 6 complex per link, SU(3) only; bad arguments are
    rejected before anything reaches the GPU."""
    from mrhs_qcu_amd import _lib

    g3 = _lib.geom((8, 4, 4, 8), 3)
    S = 4 * 4 * 4 * 8
    assert lib.qcuMrhsGauge12Bytes(g3, _lib.SINGLE) == S * 2 * 8 * 6 * 8
    assert lib.qcuMrhsGauge12Bytes(g3, _lib.HALF) * 3 == lib.qcuMrhsGaugeBytes(g3, _lib.HALF) * 2
    g2 = _lib.geom((8, 4, 4, 8), 2)
    assert lib.qcuMrhsGauge12Bytes(g2, _lib.SINGLE) == 0
    assert lib.qcuMrhsPackGauge12(16, _lib.SINGLE, 16, _lib.DOUBLE, g2, -1, None) != 0
    assert "nColors = 3" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsPackGauge12(16, _lib.SINGLE, 16, _lib.DOUBLE, g3, 0, None) != 0
    assert "t_sign" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash12(16, 16, 16, _lib.SINGLE, 2, 0, 0, 1, g2, _lib.tile(), None) != 0
    assert "nColors = 3" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash12(16, 16, 16, _lib.SINGLE, 2, 0, 0, 2, g3, _lib.tile(), None) != 0


def test_gauge8_sizes_and_validation(lib):
    """Eight-real links: 4 complex slots per link, SU(3) only, rebuilt while
    staging (tiles without LDS staging are refused); bad arguments are
    rejected before anything reaches the GPU."""
    from mrhs_qcu_amd import _lib

    g3 = _lib.geom((8, 4, 4, 8), 3)
    S = 4 * 4 * 4 * 8
    assert lib.qcuMrhsGauge8Bytes(g3, _lib.SINGLE) == S * 2 * 8 * 4 * 8
    assert lib.qcuMrhsGauge8Bytes(g3, _lib.DOUBLE) * 9 == lib.qcuMrhsGaugeBytes(g3, _lib.DOUBLE) * 4
    g2 = _lib.geom((8, 4, 4, 8), 2)
    assert lib.qcuMrhsGauge8Bytes(g2, _lib.SINGLE) == 0
    assert lib.qcuMrhsPackGauge8(16, _lib.SINGLE, 16, _lib.DOUBLE, g2, -1, None) != 0
    assert "nColors = 3" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsPackGauge8(16, _lib.SINGLE, 16, _lib.DOUBLE, g3, 0, None) != 0
    assert "t_sign" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash8(16, 16, 16, _lib.SINGLE, 2, 0, 0, 1, g2, _lib.tile(1, 128, 1), None) != 0
    assert "nColors = 3" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash8(16, 16, 16, _lib.SINGLE, 2, 0, 0, 1, g3, _lib.tile(1, 128, 0), None) != 0
    assert "link_lds" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsDslash8(16, 16, 16, _lib.SINGLE, 2, 0, 0, 2, g3, _lib.tile(1, 128, 1), None) != 0


def test_clover_sizes_and_validation(lib):
    from mrhs_qcu_amd import _lib

    g3 = _lib.geom((8, 4, 4, 8), 3)
    S = 4 * 4 * 4 * 8
    assert lib.qcuMrhsCloverBytes(g3, _lib.SINGLE) == 2 * S * 72 * 8
    assert lib.qcuMrhsCloverBytes(_lib.geom((8, 4, 4, 8), 6), _lib.SINGLE) == 0
    assert lib.qcuMrhsCloverCompute(16, None, _lib.SINGLE, 16, _lib.HALF, 0.1, g3, None) != 0
    assert "FP64 or FP32" in lib.qcuMrhsLastError().decode()
    assert lib.qcuMrhsCloverApply(16, 32, 16, _lib.SINGLE, 4, 2, g3, None) != 0
    assert "parity" in lib.qcuMrhsLastError().decode()


def test_face_sizes_and_validation(lib):
    """Face buffers: F sites of one parity x 2 spins x N colours x nrhs; a
    split dimension needs both face pointers."""
    from mrhs_qcu_amd import _lib

    L, N, k = (8, 6, 4, 10), 3, 4
    g = _lib.geom(L, N, grid=(2, 1, 1, 2))
    S = 4 * 6 * 4 * 10
    per = 2 * N * k * 8
    assert lib.qcuMrhsFaceBytes(g, k, _lib.SINGLE, 0) == (6 * 4 * 10 // 2) * per
    assert lib.qcuMrhsFaceBytes(g, k, _lib.SINGLE, 1) == S // 6 * per
    assert lib.qcuMrhsFaceBytes(g, k, _lib.SINGLE, 3) == S // 10 * per
    assert lib.qcuMrhsFaceBytes(g, k, _lib.SINGLE, 4) == 0
    # Every face holds half the sites of a slice: F = V / (2 L_mu) for each mu.
    for mu in range(4):
        assert lib.qcuMrhsFaceBytes(g, k, _lib.DOUBLE, mu) == 8 * 6 * 4 * 10 // (2 * L[mu]) * 2 * N * k * 16
    # Dimension 0 present, dimension 3 missing: rejected before any launch.
    four, keep = _lib.ptr_array([16, 0, 0, 0])
    rc = lib.qcuMrhsPackFaces(four, four, 16, 16, _lib.SINGLE, k, 0, 0, g, None)
    assert rc != 0 and "split dimension 3" in lib.qcuMrhsLastError().decode()
    null4, keep2 = _lib.ptr_array([0, 0, 0, 0])
    rc = lib.qcuMrhsPackFaces(null4, null4, 16, 16, _lib.SINGLE, k, 0, 0, g, None)
    assert rc != 0 and "split dimension 0" in lib.qcuMrhsLastError().decode()
    rc = lib.qcuMrhsDslashFaces(16, null4, null4, 16, _lib.SINGLE, k, 2, 0, g, None)
    assert rc != 0 and "parity" in lib.qcuMrhsLastError().decode()


DEATH = [
    ("lib.startDslash(0, 0)", "Qcu is not initialized."),
    ("lib.pushBackFermions(1, 1)", "Qcu is not initialized."),
    ("lib.getDslash(0, 0.0, 3, 1, 2)", "Qcu is not initialized."),
    ("lib.loadQcuGauge(1)", "Qcu is not initialized."),
    ("init(3); lib.getDslash(0, 0.0, 3, 1, 2)", "localGauge_ is nullptr, not loaded."),
    ("init(3); lib.startDslash(0, 0)", "Dslash operator not created"),
    ("init(3); lib.pushBackFermions(0, 0)", "fermion pointer is null."),
    ("init(9)", "Now we only support 2-8 colors."),
    ("init(3, L=(4, 4, 3, 4))", "lattice sizes must be positive and even."),
]


@pytest.mark.parametrize("code,message", DEATH)
def test_fatal_errors_exit_1(code, message):
    script = (
        "import ctypes, sys\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from mrhs_qcu_amd import _lib\n"
        "lib = _lib.load()\n"
        "def init(n, L=(4, 4, 4, 4)):\n"
        "    g = _lib.QcuGrid(); p = _lib.QcuParam()\n"
        "    for d in range(4):\n"
        "        g.grid_size[d] = 1; p.lattice_size[d] = L[d]\n"
        "    lib.initGridSize(ctypes.byref(g), ctypes.byref(p), n)\n"
        f"{code}\n"
        "print('returned')\n"
    )
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert message in r.stderr
    assert "returned" not in r.stdout


def test_deprecated_entry_points_are_noops(lib):
    lib.dslashQcu(None, None, None, None, 0)
    lib.fullDslashQcu(None, None, None, None, 0)


def test_fastdiv_exact(tmp_path):
    exe = tmp_path / "fastdiv_check"
    src = os.path.join(ROOT, "tests", "cpp", "fastdiv_check.cpp")
    inc = os.path.join(ROOT, "mrhs_qcu_amd", "csrc")
    subprocess.run(["hipcc", "-O2", "-std=c++17", "-I", inc, src, "-o", str(exe)], check=True, timeout=300,
                   capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout


def test_default_tiles_are_launchable_shapes(lib):
    """The tile the dispatcher uses without tuning (qcuMrhsGetTile, no GPU
    needed) always satisfies the launch rules: R divides nrhs, the block is a
    multiple of 64 up to 256, t_block / z_block divide Lt / Lz, and LDS link
    staging fits the per-workgroup budget."""
    from mrhs_qcu_amd import _lib

    lds_budget = 65536 - 2048
    bytes_per_complex = {_lib.HALF: 4, _lib.BF16: 4, _lib.SINGLE: 8, _lib.DOUBLE: 16}
    for L in ((4, 4, 4, 4), (8, 8, 8, 8), (24, 24, 24, 48), (16, 16, 16, 32), (12, 12, 12, 24), (6, 4, 10, 14)):
        for N in range(2, 9):
            for k in (1, 2, 3, 6, 12, 16, 24, 40):
                for prec in (_lib.HALF, _lib.BF16, _lib.SINGLE, _lib.DOUBLE):
                    t = _lib.Tile()
                    assert lib.qcuMrhsGetTile(_lib.geom(L, N), prec, k, ctypes.byref(t)) == 0
                    assert k % t.rhs_per_lane == 0, (L, N, k, prec, t.as_tuple())
                    assert t.block % 64 == 0 and 64 <= t.block <= 256
                    assert t.t_block >= 1 and L[3] % t.t_block == 0
                    assert t.z_block == 0 or L[2] % t.z_block == 0
                    if t.link_lds:
                        P = k // t.rhs_per_lane
                        shm = ((t.block - 1) // P + 2) * 8 * N * N * bytes_per_complex[prec]
                        assert shm <= lds_budget, (L, N, k, prec, t.as_tuple())


TUNE_CHILD = r"""
import ctypes, json, sys
sys.path.insert(0, %(root)r)
from mrhs_qcu_amd import _lib
lib = _lib.load()
out = []
for L, k in (((8, 8, 8, 8), 4), ((8, 8, 8, 8), 6), ((4, 4, 4, 8), 4)):
    t = _lib.Tile()
    assert lib.qcuMrhsGetTile(_lib.geom(L, 3), _lib.SINGLE, k, ctypes.byref(t)) == 0
    out.append(t.as_tuple())
print(json.dumps(out))
"""


def test_tune_cache_file_is_keyed_by_architecture_and_shape(tmp_path):
    """The autotuner's cache file (QCU_MRHS_TUNE_CACHE): lines for this
    process's GPU architecture ("unknown" without a GPU) and shape are used,
    later lines win, other architectures and shapes are ignored."""
    import json

    import torch

    arch = "unknown"
    if torch.cuda.is_available():
        arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    path = tmp_path / "tune.txt"
    path.write_text("gfx000 3 4 1 8 8 8 8 0 1 64 0 1 0 0 0 0\n"
                    f"{arch} 3 4 1 8 8 8 8 0 1 256 0 2 2 0 0 1\n"
                    f"{arch} 3 4 1 8 8 8 8 0 2 256 1 4 2 3 0 1\n"
                    f"{arch} 3 6 1 8 8 8 8 0 2 128 1 2 4 0 0 1\n")
    code = TUNE_CHILD % {"root": ROOT}
    env = dict(os.environ, QCU_MRHS_TUNE_CACHE=str(path))
    p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    got = json.loads(p.stdout.strip().splitlines()[-1])
    assert tuple(got[0]) == (2, 256, 1, 4, 2, 3, 0, 1)
    assert tuple(got[1]) == (2, 128, 1, 2, 4, 0, 0, 1)
    env.pop("QCU_MRHS_TUNE_CACHE")
    p = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    default = json.loads(p.stdout.strip().splitlines()[-1])
    assert tuple(got[2]) == tuple(default[2])  # shape not in the file: default tile


def test_pinned_tile_table():
    """The tile table shipped beside the library: every shape line has the
    17 fields of the tune-cache format, a launchable tile (R divides nrhs,
    block a multiple of 64 up to 256) and lists the benchmark and solver
    shapes (24^3x48 and 32^3x64, 16 RHS, SU(3), every precision, FP8 included)."""
    path = os.path.join(ROOT, "mrhs_qcu_amd", "tiles_gfx950.txt")
    shapes = set()
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        assert len(f) == 17 and f[0] == "gfx950", line
        N, k, prec, Lx, Ly, Lz, Lt, mask, R, block, lds, tb, zb, hints, occ, trav = map(int, f[1:])
        assert k % R == 0 and block in (64, 128, 192, 256) and lds in (0, 1), line
        assert trav in (0, 1, 2, 3, 4, 6), line
        if trav & 2:
            # hypercube mapping: blocking counted in 2x2x2x2 blocks, whole
            # blocks of 8 P lanes per workgroup, links staged in LDS
            P = k // R
            assert P & (P - 1) == 0 and P <= 32 and block % (8 * P) == 0 and lds == 1 and N <= 4, line
        if trav == 6:
            # column order: t_block x z_block blocks in (y, z); the columns fill
            # whole workgroups and every XCD (a multiple of 8 columns)
            by, bz = Ly // 2, Lz // 2
            assert by % tb == 0 and bz % zb == 0 and (by // tb) * (bz // zb) % 8 == 0, line
            assert (Lx // 2 * tb * zb) % (block // (8 * (k // R))) == 0, line
        elif trav == 4:
            # column order in sites
            assert Ly % tb == 0 and Lz % zb == 0 and (Ly // tb) * (Lz // zb) % 8 == 0, line
            assert (Lx // 2 * tb * zb * Lt * (k // R)) % block == 0, line
        else:
            assert Lt % max(tb, 1) == 0 and (zb == 0 or Lz % zb == 0), line
            if trav & 2:
                assert (Lt // 2) % max(tb, 1) == 0 and (zb == 0 or (Lz // 2) % zb == 0), line
        shapes.add((N, k, prec, (Lx, Ly, Lz, Lt), mask))
    for L in ((24, 24, 24, 48), (32, 32, 32, 64)):
        for prec in (0, 1, 2, 4, 5):
            assert (3, 16, prec, L, 0) in shapes, (L, prec)


def test_set_tile_round_trip(lib):
    """qcuMrhsSetTile pins a shape's tile for this process (hypercube tiles
    included, traversal bit 1); qcuMrhsGetTile returns it; malformed tiles are
    refused with a message."""
    from mrhs_qcu_amd import _lib

    g = _lib.geom((6, 4, 10, 14), 3)
    want = (2, 128, 1, 7, 5, 8, 0, 3)
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, ctypes.byref(_lib.tile(*want))) == 0
    t = _lib.Tile()
    assert lib.qcuMrhsGetTile(g, _lib.SINGLE, 8, ctypes.byref(t)) == 0
    assert t.as_tuple() == want
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, ctypes.byref(_lib.tile(3, 128))) != 0
    assert b"rhs_per_lane" in lib.qcuMrhsLastError()
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, ctypes.byref(_lib.tile(2, 96))) != 0
    assert b"block" in lib.qcuMrhsLastError()
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, ctypes.byref(_lib.tile(2, 128, 1, 1, 1, 8, 0, 3 | 4))) != 0
    assert b"traversal 4 or 6" in lib.qcuMrhsLastError()
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, ctypes.byref(_lib.tile(2, 128, 1, 1, 1, 8, 0, 2 | 8))) != 0
    assert b"traversal bits 3" in lib.qcuMrhsLastError()
    col = (2, 128, 1, 2, 5, 0, 0, 6)  # column order: 2 x 5 blocks in (y, z)
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 4, ctypes.byref(_lib.tile(*col))) == 0
    assert lib.qcuMrhsGetTile(g, _lib.SINGLE, 4, ctypes.byref(t)) == 0
    assert t.as_tuple() == col
    assert lib.qcuMrhsSetTile(g, _lib.SINGLE, 8, None) != 0
