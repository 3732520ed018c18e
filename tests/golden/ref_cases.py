"""Cases of the reference-kernel golden vectors (inputs are regenerated from
the seeds; their SHA-256 is stored beside the outputs)."""
import hashlib

import numpy as np

from mrhs_qcu_amd import fields

# name, L, N, nrhs, prec (2 = FP64, 1 = FP32), parity, dagger, grid, seed
CASES = [
    ("n3_k1_p0_d0", (4, 4, 4, 4), 3, 1, 2, 0, 0, (1, 1, 1, 1), 100),
    ("n3_k2_p1_d1", (4, 4, 4, 4), 3, 2, 2, 1, 1, (1, 1, 1, 1), 101),
    ("n3_k5_p0_d1_gridt", (4, 4, 4, 4), 3, 5, 2, 0, 1, (1, 1, 1, 2), 102),
    ("n4_k3_p1_d0_grid2222", (4, 4, 4, 4), 4, 3, 2, 1, 0, (2, 2, 2, 2), 103),
    ("n5_k1_p0_d0", (4, 4, 4, 8), 5, 1, 2, 0, 0, (1, 1, 1, 1), 104),
    ("n3_k2_p1_d0_fp32", (4, 4, 4, 4), 3, 2, 1, 1, 0, (1, 1, 1, 1), 105),
    ("n3_k1_p1_d1_8x8x8x8", (8, 8, 8, 8), 3, 1, 2, 1, 1, (1, 1, 1, 1), 0),
]


def inputs(case):
    name, L, N, k, prec, parity, dagger, grid, seed = case
    g = fields.gauge_field(L, N, seed=seed)
    psi = fields.gaussian_spinors(L, N, k, seed=seed + 1000)
    if prec == 1:
        g = g.astype(np.complex64)
        psi = psi.astype(np.complex64)
    return g, psi


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
