"""The autotuner's persistent cache (QCU_MRHS_TUNE_CACHE): a tile tuned in
one process is written to the file and returned by qcuMrhsGetTile in a fresh
process without tuning again. Runs two short child processes, one after the
other."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import ctypes, json, sys
sys.path.insert(0, %(root)r)
import torch
from mrhs_qcu_amd import _lib
lib = _lib.load()
L, N, k, prec = (8, 8, 8, 8), 3, 4, _lib.SINGLE
g = _lib.geom(L, N)
S = 4 * 8 * 8 * 8
st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
if sys.argv[1] == "tune":
    gauge = torch.randn(S * 2 * 8 * N * N * 2, device="cuda:0")
    a = torch.randn(S * 4 * N * k * 2, device="cuda:0")
    o = torch.empty_like(a)
    best = _lib.Tile()
    us = ctypes.c_double()
    _lib.check(lib.qcuMrhsTune(o.data_ptr(), a.data_ptr(), gauge.data_ptr(), prec, k, g, ctypes.byref(best),
                               ctypes.byref(us), st))
    print(json.dumps(best.as_tuple()))
else:
    t = _lib.Tile()
    _lib.check(lib.qcuMrhsGetTile(g, prec, k, ctypes.byref(t)))
    print(json.dumps(t.as_tuple()))
"""


def run_child(mode, env):
    code = CHILD % {"root": ROOT}
    p = subprocess.run([sys.executable, "-c", code, mode], env=env, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    return tuple(json.loads(p.stdout.strip().splitlines()[-1]))


def test_tune_cache_file_round_trip(tmp_path):
    path = tmp_path / "tune.txt"
    env = dict(os.environ, QCU_MRHS_TUNE_CACHE=str(path))
    tuned = run_child("tune", env)
    assert path.exists() and path.read_text().strip(), "no cache line written"
    assert run_child("get", env) == tuned
    # Without the file a fresh process falls back to its default tile.
    env.pop("QCU_MRHS_TUNE_CACHE")
    default = run_child("get", env)
    assert len(default) == len(tuned)


PINNED = r"""
import ctypes, json, sys
sys.path.insert(0, %(root)r)
import torch
from mrhs_qcu_amd import _lib
lib = _lib.load()
torch.zeros(1, device="cuda:0")
t = _lib.Tile()
_lib.check(lib.qcuMrhsGetTile(_lib.geom((24, 24, 24, 48), 3), _lib.SINGLE, 16, ctypes.byref(t)))
print(json.dumps(t.as_tuple()))
"""


def test_pinned_tiles_used_without_tuning():
    """A fresh process without tuning gets the headline shape's tile from the
    table shipped beside the library (tiles_gfx950.txt)."""
    want = None
    for line in open(os.path.join(ROOT, "mrhs_qcu_amd", "tiles_gfx950.txt")):
        f = line.split()
        if len(f) == 17 and f[0] == "gfx950" and f[1:9] == ["3", "16", "1", "24", "24", "24", "48", "0"]:
            want = tuple(int(v) for v in f[9:])
    assert want is not None
    env = dict(os.environ)
    env.pop("QCU_MRHS_TILES", None)
    env.pop("QCU_MRHS_TUNE_CACHE", None)
    p = subprocess.run([sys.executable, "-c", PINNED % {"root": ROOT}], env=env, capture_output=True, text=True,
                       timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    assert tuple(json.loads(p.stdout.strip().splitlines()[-1])) == want
