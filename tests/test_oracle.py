"""Pins the CPU oracle (oracle/oracle.c) without a GPU.

Each check is independent of the oracle's per-direction restatement:
  * the Dirac matrices satisfy the Clifford algebra and g5 = diag(1,1,-1,-1);
  * the oracle equals the textbook operator built from full coordinates and
    the DeGrand-Rossi matrices (a dropped term, a wrong sign, a swapped spin
    or a transposed link fails it), for N = 2..5, k = 1..3, both parities,
    both dagger flags, periodic and split (grid > 1) dimensions;
  * free field: plane waves are eigenvectors with the closed-form symbol,
    including the antiperiodic time boundary folded into the links;
  * g5-hermiticity, gauge covariance, linearity and RHS independence.
"""
import numpy as np
import pytest

from mrhs_qcu_amd import fields
from oracle import oracle
from helpers import GAMMA, GAMMA5, textbook_dslash, rel_err, cb_index


def test_clifford_algebra():
    for m in range(4):
        assert np.allclose(GAMMA[m], GAMMA[m].conj().T)
        for n in range(4):
            anti = GAMMA[m] @ GAMMA[n] + GAMMA[n] @ GAMMA[m]
            assert np.allclose(anti, 2.0 * (m == n) * np.eye(4))
    assert np.allclose(GAMMA5, np.diag([1, 1, -1, -1]))


@pytest.mark.parametrize("N", [2, 3, 4, 5])
@pytest.mark.parametrize("parity", [0, 1])
@pytest.mark.parametrize("dagger", [0, 1])
def test_oracle_matches_textbook(N, parity, dagger):
    L = (4, 4, 2, 6) if N != 3 else (4, 4, 4, 4)
    g = fields.gauge_field(L, N, seed=10 + N)
    k = 1 + (N % 3)
    psi = fields.gaussian_spinors(L, N, k, seed=20 + N)
    got = oracle.dslash(g, psi, L, parity, dagger)
    want = textbook_dslash(g, psi, L, parity, dagger)
    assert rel_err(got, want) < 1e-13


@pytest.mark.parametrize("grid", [(2, 1, 1, 1), (1, 2, 1, 1), (1, 1, 2, 1), (1, 1, 1, 2), (2, 2, 2, 2)])
@pytest.mark.parametrize("parity", [0, 1])
def test_oracle_boundary_skip(grid, parity):
    L = (4, 4, 4, 4)
    g = fields.gauge_field(L, 3, seed=31)
    psi = fields.gaussian_spinors(L, 3, 2, seed=32)
    got = oracle.dslash(g, psi, L, parity, 1, grid=grid)
    want = textbook_dslash(g, psi, L, parity, 1, grid=grid)
    assert rel_err(got, want) < 1e-13
    # and it differs from the periodic operator
    per = oracle.dslash(g, psi, L, parity, 1)
    assert rel_err(got, per) > 1e-3


@pytest.mark.parametrize("antiperiodic", [False, True])
@pytest.mark.parametrize("dagger", [0, 1])
def test_free_field_plane_wave(antiperiodic, dagger):
    """U = 1 (with -1 on the last time slice when antiperiodic): psi(x) =
    exp(i p.x) u is mapped to exp(i p.x) [sum 2 cos p - 2 i s sum g sin p] u."""
    L = (6, 4, 4, 8)
    N = 3
    S = fields.half_volume(L)
    g = np.zeros((4, 2, S, N, N), dtype=np.complex128)
    g[...] = np.eye(N)
    if antiperiodic:
        g = fields.fold_antiperiodic_t(g, L)
    n = (1, 2, 1, 3)
    p = [2 * np.pi * n[m] / L[m] for m in range(4)]
    if antiperiodic:
        p[3] = (2 * n[3] + 1) * np.pi / L[3]
    rng = np.random.default_rng(5)
    u = rng.standard_normal((4, N)) + 1j * rng.standard_normal((4, N))
    s = -1.0 if dagger else 1.0
    sym = sum(2 * np.cos(p[m]) * np.eye(4) - 2j * s * np.sin(p[m]) * GAMMA[m] for m in range(4))
    for parity in (0, 1):
        xh, y, z, t = fields.site_coords(L)
        Xin = 2 * xh + ((1 - parity + y + z + t) & 1)
        ph_in = np.exp(1j * (p[0] * Xin + p[1] * y + p[2] * z + p[3] * t))
        psi = (ph_in[:, None, None] * u[None])[None]
        Xout = 2 * xh + ((parity + y + z + t) & 1)
        ph_out = np.exp(1j * (p[0] * Xout + p[1] * y + p[2] * z + p[3] * t))
        want = (ph_out[:, None, None] * (sym @ u)[None])[None]
        got = oracle.dslash(g, psi, L, parity, dagger)
        assert rel_err(got, want) < 1e-13


def test_constant_field_gives_eight():
    L = (4, 4, 4, 4)
    S = fields.half_volume(L)
    g = np.zeros((4, 2, S, 3, 3), dtype=np.complex128)
    g[...] = np.eye(3)
    psi = np.ones((1, S, 4, 3), dtype=np.complex128) * (0.3 - 0.7j)
    got = oracle.dslash(g, psi, L, 0, 0)
    assert np.allclose(got, 8 * psi, atol=1e-14)


@pytest.mark.parametrize("N", [2, 3, 5])
def test_gamma5_hermiticity(N):
    """<chi, D(s=+1)_{p<-1-p} psi> = <D(s=-1)_{1-p<-p} chi, psi>."""
    L = (4, 4, 4, 6)
    g = fields.gauge_field(L, N, seed=40 + N)
    for p in (0, 1):
        psi = fields.gaussian_spinors(L, N, 1, seed=41)
        chi = fields.gaussian_spinors(L, N, 1, seed=42)
        lhs = np.vdot(chi, oracle.dslash(g, psi, L, p, 0))
        rhs = np.vdot(oracle.dslash(g, chi, L, 1 - p, 1), psi)
        assert abs(lhs - rhs) < 1e-10 * abs(lhs)
        # dagger flag = g5 D g5
        g5psi = np.einsum("ab,ksbc->ksac", GAMMA5, psi)
        a = oracle.dslash(g, psi, L, p, 1)
        b = np.einsum("ab,ksbc->ksac", GAMMA5, oracle.dslash(g, g5psi, L, p, 0))
        assert rel_err(a, b) < 1e-13


def test_gauge_covariance():
    L = (4, 4, 4, 4)
    N = 3
    S = fields.half_volume(L)
    g = fields.gauge_field(L, N, seed=50)
    rng = np.random.Generator(np.random.PCG64(51))
    om = fields.random_sun((2, S), N, rng)  # per parity, per site
    xh, y, z, t = fields.site_coords(L)
    gp = np.empty_like(g)
    for p in range(2):
        X = 2 * xh + ((p + y + z + t) & 1)
        for mu in range(4):
            c = [X.copy(), y.copy(), z.copy(), t.copy()]
            c[mu] = (c[mu] + 1) % L[mu]
            nidx = cb_index(L, *c)
            gp[mu, p] = om[p] @ g[mu, p] @ np.conj(np.swapaxes(om[1 - p, nidx], -1, -2))
    for p in (0, 1):
        psi = fields.gaussian_spinors(L, N, 2, seed=52)
        psi_t = np.einsum("sij,ksaj->ksai", om[1 - p], psi)
        got = oracle.dslash(gp, psi_t, L, p, 0)
        want = np.einsum("sij,ksaj->ksai", om[p], oracle.dslash(g, psi, L, p, 0))
        assert rel_err(got, want) < 1e-13


def test_linearity_and_rhs_independence():
    L = (4, 4, 4, 4)
    g = fields.gauge_field(L, 3, seed=60)
    psi = fields.gaussian_spinors(L, 3, 9, seed=61)
    many = oracle.dslash(g, psi, L, 1, 0)
    for r in range(9):
        one = oracle.dslash(g, psi[r:r + 1], L, 1, 0)
        assert np.array_equal(one[0], many[r])
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    comb = oracle.dslash(g, a * psi[:1] + b * psi[1:2], L, 1, 0)
    assert rel_err(comb, a * many[:1] + b * many[1:2]) < 1e-13


def test_single_precision_oracle():
    L = (4, 4, 4, 4)
    g = fields.gauge_field(L, 3, seed=70)
    psi = fields.gaussian_spinors(L, 3, 3, seed=71)
    d64 = oracle.dslash(g, psi, L, 0, 0)
    d32 = oracle.dslash(g, psi, L, 0, 0, single=True)
    assert rel_err(d32, d64) < 1e-6


def test_shift_index_maps_roundtrip():
    L = (4, 4, 2, 2)
    S, N = fields.half_volume(L), 3
    v = np.arange(S * 4 * N, dtype=np.float64) + 1j * np.arange(S * 4 * N)
    soa = oracle.shift_vector(v, True, S, N)
    assert soa[1] == v[4 * N]  # element e = 0 of site 1
    assert np.array_equal(oracle.shift_vector(soa, False, S, N), v)
    gg = np.arange(8 * S * N * N) * (1 + 1j)
    gs = oracle.shift_gauge(gg, True, S, N)
    assert np.array_equal(oracle.shift_gauge(gs, False, S, N), gg)
    assert np.array_equal(np.sort(gs.real), np.sort(gg.real))


def test_generators():
    L = (4, 4, 4, 4)
    for N in (2, 3, 4, 5):
        g = fields.gauge_field(L, N, seed=1, antiperiodic_t=False)
        u = g.reshape(-1, N, N)
        eye = np.einsum("bij,bkj->bik", u, np.conj(u))
        assert np.allclose(eye, np.eye(N), atol=1e-12)
        assert np.allclose(np.linalg.det(u), 1.0, atol=1e-12)
    gs = fields.gauge_field(L, 3, seed=1, smear_eps=0.3, antiperiodic_t=False).reshape(-1, 3, 3)
    g0 = fields.gauge_field(L, 3, seed=1, antiperiodic_t=False).reshape(-1, 3, 3)
    assert np.allclose(np.linalg.det(gs), 1.0, atol=1e-12)
    # smearing moves links toward the identity
    d_s = np.mean(np.abs(gs - np.eye(3)))
    d_0 = np.mean(np.abs(g0 - np.eye(3)))
    assert d_s < d_0
    # antiperiodic fold flips exactly the last time slice of U_t
    ga = fields.gauge_field(L, 3, seed=1, antiperiodic_t=True)
    gp = fields.gauge_field(L, 3, seed=1, antiperiodic_t=False)
    t = fields.site_coords(L)[3]
    assert np.array_equal(ga[3][:, t == 3], -gp[3][:, t == 3])
    assert np.array_equal(ga[3][:, t < 3], gp[3][:, t < 3])
    assert np.array_equal(ga[:3], gp[:3])
