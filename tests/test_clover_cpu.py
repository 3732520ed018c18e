"""The textbook clover term of the tests (helpers.textbook_clover), pinned by
properties before it serves as the reference of the GPU clover kernels:
sigma algebra in the DeGrand-Rossi basis (block diagonal, since gamma_5 is),
free field A = 1, Hermiticity, gauge covariance A -> g A g^+, the abelian
closed form for a constant U(1) field, and linearity in coeff."""
import numpy as np
import pytest

from mrhs_qcu_amd import fields
from helpers import GAMMA, GAMMA5, sigma, textbook_clover, clover_blocks, gauge_transform


def test_sigma_algebra():
    assert np.allclose(GAMMA5, np.diag([1, 1, -1, -1]))
    for mu in range(4):
        for nu in range(4):
            s = sigma(mu, nu)
            assert np.allclose(s, s.conj().T)
            assert np.allclose(s, -sigma(nu, mu))
            assert np.allclose(s[:2, 2:], 0) and np.allclose(s[2:, :2], 0)
            assert np.allclose(s @ GAMMA5, GAMMA5 @ s)


def test_free_field_is_identity():
    L, N = (4, 4, 4, 4), 3
    g = np.tile(np.eye(N, dtype=np.complex128), (4, 2, fields.half_volume(L), 1, 1))
    A = textbook_clover(g, L, 0.9)
    assert np.allclose(A, np.eye(4 * N))


def test_hermitian_and_block_diagonal():
    L, N = (4, 4, 4, 4), 3
    A = textbook_clover(fields.gauge_field(L, N, seed=1), L, 0.7)
    assert np.abs(A - np.conj(np.swapaxes(A, -1, -2))).max() < 1e-14
    B = 2 * N
    assert np.abs(A[..., :B, B:]).max() < 1e-15 and np.abs(A[..., B:, :B]).max() < 1e-15
    assert clover_blocks(A, N).shape == (2, fields.half_volume(L), 2, B, B)


def test_gauge_covariance():
    L, N = (4, 4, 4, 4), 3
    S = fields.half_volume(L)
    g = fields.gauge_field(L, N, seed=2)
    rng = np.random.Generator(np.random.PCG64(3))
    om = fields.random_sun((2, S), N, rng)
    A = textbook_clover(g, L, 0.5)
    At = textbook_clover(gauge_transform(g, L, om), L, 0.5)
    G = np.einsum("ab,psij->psaibj", np.eye(4), om).reshape(2, S, 4 * N, 4 * N)
    want = G @ A @ np.conj(np.swapaxes(G, -1, -2))
    assert np.abs(At - want).max() < 1e-13


def test_constant_abelian_field_closed_form():
    """U(1) links U_mu(x) = exp(i theta_mu(x)) with a constant field strength
    in the (x, y) plane: every plaquette, hence every leaf, is exp(i phi), so
    Q = 4 exp(i phi), F_xy = sin(phi) (times the identity in colour: the U(1)
    field embedded as multiples of the N x N identity) and
    A = 1 + coeff sigma_xy sin(phi)."""
    L, N = (4, 4, 4, 4), 2
    phi = 2 * np.pi / (L[0] * L[1]) * 3  # flux quantised on the torus
    S = fields.half_volume(L)
    xh, y, z, t = fields.site_coords(L)
    g = np.tile(np.eye(N, dtype=np.complex128), (4, 2, S, 1, 1))
    for p in range(2):
        X = 2 * xh + ((p + y + z + t) & 1)
        # Landau-type gauge: U_y(x) = exp(i phi x); the wrap in x is fixed by U_x at x = Lx - 1.
        uy = np.exp(1j * phi * X)
        ux = np.where(X == L[0] - 1, np.exp(-1j * phi * L[0] * y), 1.0)
        g[1, p] = uy[:, None, None] * np.eye(N)
        g[0, p] = ux[:, None, None] * np.eye(N)
    coeff = 0.8
    A = textbook_clover(g, L, coeff)
    want = np.eye(4 * N) + coeff * np.sin(phi) * np.kron(sigma(0, 1), np.eye(N))
    assert np.abs(A - want).max() < 1e-13


def test_linear_in_coeff():
    L, N = (4, 4, 4, 4), 2
    g = fields.gauge_field(L, N, seed=4)
    A1, A2 = textbook_clover(g, L, 1.0), textbook_clover(g, L, 2.0)
    I = np.eye(4 * N)
    assert np.abs((A2 - I) - 2 * (A1 - I)).max() < 1e-14
