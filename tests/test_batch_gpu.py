"""Batched hopping term (qcuMrhsBatchCreate / Run / Destroy): uneven
problems (SU(2)..SU(5), different lattices and nrhs, FP32) applied by one
HIP graph (a single work-list kernel; with an N outside 2..5 a chain of
kernels), checked against the FP64 oracle and against the same kernels with
the same tiles launched one by one, on several replays with eager work in
between."""
import ctypes

import numpy as np
import pytest
import torch

from mrhs_qcu_amd import _lib, fields
from oracle import oracle
from helpers import rel_err
from test_dslash_gpu import pack_inputs, unpack, stream

pytestmark = pytest.mark.gpu

# (L, N, nrhs, parity, dagger): the shapes of the uneven workload at sizes the
# oracle checks in seconds.
PROBLEMS = [((8, 8, 8, 8), 2, 8, 0, 0), ((8, 8, 8, 16), 3, 12, 1, 0), ((8, 4, 4, 8), 4, 24, 0, 1),
            ((4, 4, 4, 8), 5, 6, 1, 1)]


# The same colour counts on other shapes, with column-order tiles set for them
# (traversal 4: columns of cy x cz sites in (y, z), XCD x walking columns
# 8 g + x along t), in the work list's tile family (R = 2 with links in LDS
# for N <= 4, R = 1 for N = 5; 128 lanes per workgroup).
COLUMN_PROBLEMS = [((8, 8, 16, 8), 2, 8, 1, 0), ((8, 16, 8, 8), 3, 12, 0, 1), ((8, 8, 8, 8), 4, 24, 1, 1),
                   ((4, 8, 8, 4), 5, 6, 0, 0)]
COLUMN_TILES = [(2, 128, 1, 2, 4, 0, 0, 4), (2, 128, 1, 4, 2, 0, 0, 4), (2, 128, 1, 2, 2, 0, 0, 4),
                (1, 128, 0, 4, 2, 0, 0, 4)]


@pytest.mark.parametrize("problems", [PROBLEMS, [((4, 4, 4, 8), 6, 4, 1, 0), ((8, 4, 4, 8), 3, 10, 0, 1)],
                                      COLUMN_PROBLEMS],
                         ids=["worklist", "chain", "worklist_columns"])
def test_batch_graph_replays(lib, dev, problems):
    if problems is COLUMN_PROBLEMS:
        for (L, N, k, _, _), t in zip(problems, COLUMN_TILES):
            _lib.check(lib.qcuMrhsSetTile(_lib.geom(L, N), _lib.SINGLE, k, ctypes.byref(_lib.tile(*t))))
    items = []
    for i, (L, N, k, parity, dagger) in enumerate(problems):
        g = fields.gauge_field(L, N, seed=40 + i)
        psi = fields.gaussian_spinors(L, N, k, seed=50 + i)
        geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
        pout = torch.zeros_like(pin)
        want = oracle.dslash(g, psi, L, parity, dagger)
        items.append(dict(L=L, N=N, k=k, parity=parity, dagger=dagger, geom=geom, gnat=gnat, pin=pin, pout=pout,
                          want=want, shape=psi.shape[1:]))
    probs = (_lib.Problem * len(items))()
    for p, it in zip(probs, items):
        p.geom = it["geom"]
        p.nrhs, p.parity, p.dagger = it["k"], it["parity"], it["dagger"]
        p.out, p.in_, p.gauge = it["pout"].data_ptr(), it["pin"].data_ptr(), it["gnat"].data_ptr()
    handle = ctypes.c_void_p()
    _lib.check(lib.qcuMrhsBatchCreate(probs, len(items), ctypes.byref(handle), stream()))
    try:
        # The same kernels launched one by one with the batch's (tuned) tiles.
        eager = []
        for i, it in enumerate(items):
            t = _lib.Tile()  # the tile the batch runs this problem with
            _lib.check(lib.qcuMrhsBatchGetTile(handle, i, ctypes.byref(t)))
            if problems is COLUMN_PROBLEMS:
                assert t.as_tuple() == COLUMN_TILES[i], (i, t.as_tuple())
            o = torch.empty_like(it["pin"])
            _lib.check(lib.qcuMrhsDslash(o.data_ptr(), it["pin"].data_ptr(), it["gnat"].data_ptr(), _lib.SINGLE,
                                         it["k"], it["parity"], it["dagger"], it["geom"], ctypes.byref(t), stream()))
            eager.append(o)
        torch.cuda.synchronize()
        for replay in range(3):
            for it in items:
                it["pout"].fill_(float("nan"))
            scratch = torch.randn(1 << 20, device=dev)
            scratch = (scratch * 2.0).sum()  # unrelated eager work between replays
            _lib.check(lib.qcuMrhsBatchRun(handle, stream()))
            torch.cuda.synchronize()
            for it, o in zip(items, eager):
                assert torch.equal(it["pout"], o), (replay, it["N"])
                got = unpack(lib, dev, it["pout"], _lib.SINGLE, it["k"], it["geom"], it["shape"])
                err = rel_err(got, it["want"])
                assert err < 1e-6, (replay, it["N"], err)
    finally:
        _lib.check(lib.qcuMrhsBatchDestroy(handle))


def test_batch_rejects_bad_problem(lib, dev):
    L, N, k = (4, 4, 4, 4), 3, 2
    g = fields.gauge_field(L, N, seed=1)
    psi = fields.gaussian_spinors(L, N, k, seed=2)
    geom, gnat, pin = pack_inputs(lib, dev, g, psi, L, N, _lib.SINGLE)
    pout = torch.empty_like(pin)
    probs = (_lib.Problem * 1)()
    probs[0].geom = geom
    probs[0].nrhs, probs[0].parity, probs[0].dagger = k, 0, 0
    probs[0].out, probs[0].in_, probs[0].gauge = pout.data_ptr(), pin.data_ptr(), None
    handle = ctypes.c_void_p()
    assert lib.qcuMrhsBatchCreate(probs, 1, ctypes.byref(handle), stream()) != 0
    assert "null field" in lib.qcuMrhsLastError().decode()
    assert not handle.value
