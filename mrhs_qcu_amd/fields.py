"""Synthetic lattice fields for tests and benchmarks (numpy, host side).

Conventions (the qcu.h user layouts):
  * a single-parity spinor is ``[S][4][N]`` complex128, S = V/2 checkerboard
    sites, checkerboard index ``idx = ((t*Lz + z)*Ly + y)*(Lx/2) + x//2``; a set
    of k right-hand sides is ``[k][S][4][N]``;
  * the gauge field is ``[4][2][S][N][N]`` complex128 (direction, parity,
    site, row, column), row-major links.

Generators are seeded (``numpy.random.PCG64``) so every machine with the same
numpy produces the same fields:
  * links: complex Gaussian N x N, rows orthonormalised by modified
    Gram-Schmidt, last row multiplied by conj(det) so det = 1 (SU(N));
  * smeared links: V -> V**eps through the principal branch of log V
    (V**eps = W diag(exp(i eps theta)) W^-1), re-projected to SU(N): a
    geodesic step toward the identity that mimics a thermalised field;
  * antiperiodic time: U_t(x) -> -U_t(x) on the last time slice, applied to
    the links so the periodic operator realises the antiperiodic boundary;
  * spinors: every component Re, Im ~ N(0, 1); point sources: a unit delta in
    one spin-colour component at a seeded site.
"""
import numpy as np


def half_volume(L):
    return (L[0] // 2) * L[1] * L[2] * L[3]


def site_coords(L):
    """(xh, y, z, t) arrays of every checkerboard index."""
    subLx, Ly, Lz, Lt = L[0] // 2, L[1], L[2], L[3]
    idx = np.arange(half_volume(L))
    xh = idx % subLx
    y = (idx // subLx) % Ly
    z = (idx // (subLx * Ly)) % Lz
    t = idx // (subLx * Ly * Lz)
    return xh, y, z, t


def full_x(L, parity):
    """Full x coordinate of every checkerboard index of a parity."""
    xh, y, z, t = site_coords(L)
    return 2 * xh + ((parity + y + z + t) & 1)


def project_sun(m):
    """Modified Gram-Schmidt on the rows of a batch of N x N matrices, then
    fix the determinant to 1 by rescaling the last row."""
    m = np.array(m, dtype=np.complex128, copy=True)
    n = m.shape[-1]
    for i in range(n):
        for j in range(i):
            proj = np.sum(np.conj(m[..., j, :]) * m[..., i, :], axis=-1, keepdims=True)
            m[..., i, :] -= proj * m[..., j, :]
        nrm = np.sqrt(np.sum(np.abs(m[..., i, :]) ** 2, axis=-1, keepdims=True))
        m[..., i, :] /= nrm
    det = np.linalg.det(m)
    m[..., n - 1, :] *= np.conj(det)[..., None]
    return m


def random_sun(shape, N, rng):
    g = rng.standard_normal(shape + (N, N)) + 1j * rng.standard_normal(shape + (N, N))
    return project_sun(g)


def smear_toward_identity(u, eps):
    """u**eps on the principal branch, re-projected to SU(N)."""
    w, v = np.linalg.eig(u)
    theta = np.angle(w)
    d = np.exp(1j * eps * theta)
    out = (v * d[..., None, :]) @ np.linalg.inv(v)
    return project_sun(out)


def fold_antiperiodic_t(gauge, L):
    """Multiply U_t on the last time slice by -1 (both parities)."""
    g = gauge.copy()
    t = site_coords(L)[3]
    last = t == L[3] - 1
    g[3, :, last] *= -1.0
    return g


def gauge_field(L, N, seed=0, smear_eps=None, antiperiodic_t=True, chunk=1 << 20):
    """User-layout gauge [4][2][S][N][N] complex128."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = half_volume(L)
    total = 4 * 2 * S
    out = np.empty((total, N, N), dtype=np.complex128)
    for a in range(0, total, chunk):
        b = min(total, a + chunk)
        u = random_sun((b - a,), N, rng)
        if smear_eps is not None:
            u = smear_toward_identity(u, smear_eps)
        out[a:b] = u
    out = out.reshape(4, 2, S, N, N)
    if antiperiodic_t:
        out = fold_antiperiodic_t(out, L)
    return out


def gaussian_spinors(L, N, k, seed=1, dtype=np.complex128):
    """k user-layout spinors [k][S][4][N]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = half_volume(L)
    shape = (k, S, 4, N)
    re = rng.standard_normal(shape)
    im = rng.standard_normal(shape)
    return (re + 1j * im).astype(dtype)


def point_sources(L, N, k, seed=2):
    """k spinors each a unit delta at a seeded site in spin-colour component
    r % (4N) (r = RHS index)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S = half_volume(L)
    out = np.zeros((k, S, 4, N), dtype=np.complex128)
    sites = rng.integers(0, S, size=k)
    for r in range(k):
        sc = r % (4 * N)
        out[r, sites[r], sc // N, sc % N] = 1.0
    return out


def cg_sources(L, N, k, seed=3, n_point=None):
    """Config of the CG benchmark: RHS 0..n_point-1 point sources, the rest
    Gaussian noise (n_point defaults to 4N, capped at k)."""
    n_point = min(k, 4 * N) if n_point is None else min(n_point, k)
    out = np.empty((k, half_volume(L), 4, N), dtype=np.complex128)
    if n_point:
        out[:n_point] = point_sources(L, N, n_point, seed=seed)
    if k > n_point:
        out[n_point:] = gaussian_spinors(L, N, k - n_point, seed=seed + 1)
    return out


def to_packed(psi):
    """[k][S][4][N] -> packed [S][4][N][k] (RHS innermost)."""
    return np.ascontiguousarray(np.transpose(psi, (1, 2, 3, 0)))


def from_packed(p):
    """packed [S][4][N][k] -> [k][S][4][N]."""
    return np.ascontiguousarray(np.transpose(p, (3, 0, 1, 2)))


def gauge_field_device(L, N, seed=0, device="cuda:0", antiperiodic_t=True, dtype=None):
    """The same construction as gauge_field (complex Gaussian, modified
    Gram-Schmidt on rows, det fixed to 1, antiperiodic time folded in), done
    with torch on a device for the large benchmark lattices. Returns a
    complex128 tensor [4][2][S][N][N]; the random stream is torch's, so the
    field differs from gauge_field's numpy stream with the same seed."""
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    S = half_volume(L)
    shape = (4 * 2 * S, N, N)
    m = torch.complex(torch.randn(shape, generator=gen, device=device, dtype=torch.float64),
                      torch.randn(shape, generator=gen, device=device, dtype=torch.float64))
    for i in range(N):
        for j in range(i):
            proj = torch.sum(torch.conj(m[:, j, :]) * m[:, i, :], dim=-1, keepdim=True)
            m[:, i, :] -= proj * m[:, j, :]
        m[:, i, :] /= torch.linalg.vector_norm(m[:, i, :], dim=-1, keepdim=True)
    det = torch.linalg.det(m)
    m[:, N - 1, :] *= torch.conj(det)[:, None]
    m = m.reshape(4, 2, S, N, N)
    if antiperiodic_t:
        t = torch.as_tensor(site_coords(L)[3], device=device)
        m[3, :, t == L[3] - 1] *= -1.0
    return m
