"""The oracle against the REFERENCE's own outputs (CPU).

tests/golden/ref_golden.npz holds the outputs of the reference's dslash and
shift kernels, built from the reference sources (oracle/Makefile, target
`ref`) and run on an MI355X by tests/golden/make_ref_golden.py, for the seeded
inputs of tests/golden/ref_cases.py (their SHA-256 is stored and re-checked,
which also pins the input generators).
"""
import os

import numpy as np
import pytest

from mrhs_qcu_amd import fields
from oracle import oracle
from helpers import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "ref_golden.npz")

import sys  # noqa: E402

sys.path.insert(0, os.path.join(HERE, "golden"))
import ref_cases  # noqa: E402


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("case", ref_cases.CASES, ids=[c[0] for c in ref_cases.CASES])
def test_oracle_reproduces_reference_kernel(golden, case):
    name, L, N, k, prec, parity, dagger, grid, seed = case
    g, psi = ref_cases.inputs(case)
    assert str(golden[name + "__sha"]) == ref_cases.digest(g, psi), "input generators changed"
    want = golden[name + "__out"]
    got = oracle.dslash(g, psi, L, parity, dagger, grid=grid, single=(prec == 1))
    tol = 1e-13 if prec == 2 else 1e-6
    assert rel_err(got, want) < tol


def test_oracle_shift_maps_reproduce_reference_kernels(golden):
    L, N = (4, 4, 4, 4), 3
    S = fields.half_volume(L)
    v = fields.gaussian_spinors(L, N, 1, seed=7)[0].reshape(-1)
    g = fields.gauge_field(L, N, seed=8).reshape(-1)
    assert str(golden["shift_sha"]) == ref_cases.digest(v, g)
    assert np.array_equal(oracle.shift_vector(v, True, S, N), golden["shift_vector_soa"])
    assert np.array_equal(oracle.shift_gauge(g, True, S, N), golden["shift_gauge_soa"])
