"""Test helpers: the Dirac matrices, a textbook construction of the hopping
term from full lattice coordinates, the native-layout restatements and GPU
buffer plumbing."""
import numpy as np

from mrhs_qcu_amd import fields

# DeGrand-Rossi basis.
GAMMA = np.array([
    [[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]],
], dtype=np.complex128)
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]


def cb_index(L, x, y, z, t):
    """Checkerboard index of full coordinates."""
    return ((t * L[2] + z) * L[1] + y) * (L[0] // 2) + x // 2


def textbook_dslash(gauge, psi, L, parity, dagger, grid=(1, 1, 1, 1)):
    """sum_mu U_mu(x)(1 - s g_mu) psi(x+mu) + U_mu(x-mu)^+ (1 + s g_mu) psi(x-mu),
    built from full coordinates, periodic, terms crossing the boundary of a
    dimension with grid > 1 removed. psi: [k][S][4][N] on parity 1-parity."""
    s = -1.0 if dagger else 1.0
    xh, y, z, t = fields.site_coords(L)
    X = 2 * xh + ((parity + y + z + t) & 1)
    coords = [X, y, z, t]
    idx = np.arange(fields.half_volume(L))
    out = np.zeros_like(psi)
    eye = np.eye(4)
    for mu in range(4):
        for fwd in (True, False):
            c = [a.copy() for a in coords]
            c[mu] = (c[mu] + (1 if fwd else -1)) % L[mu]
            nidx = cb_index(L, *c)
            if fwd:
                U = gauge[mu, parity, idx]
                P = eye - s * GAMMA[mu]
            else:
                U = np.conj(np.swapaxes(gauge[mu, 1 - parity, nidx], -1, -2))
                P = eye + s * GAMMA[mu]
            term = np.einsum("ab,ksbc->ksac", P, psi[:, nidx])
            term = np.einsum("sij,ksaj->ksai", U, term)
            if grid[mu] > 1:
                cross = coords[mu] == (L[mu] - 1 if fwd else 0)
                term[:, cross] = 0
            out += term
    return out


def native_gauge(gauge, L):
    """User gauge [4][2][S][N][N] -> native [2][S][8][N][N]: slots 0..3 U_mu(x),
    slots 4..7 U_mu(x - mu)^+ (restated from full coordinates)."""
    S = fields.half_volume(L)
    N = gauge.shape[-1]
    out = np.empty((2, S, 8, N, N), dtype=gauge.dtype)
    xh, y, z, t = fields.site_coords(L)
    for p in range(2):
        X = 2 * xh + ((p + y + z + t) & 1)
        coords = [X, y, z, t]
        for mu in range(4):
            out[p, :, mu] = gauge[mu, p]
            c = [a.copy() for a in coords]
            c[mu] = (c[mu] - 1) % L[mu]
            nidx = cb_index(L, *c)
            out[p, :, 4 + mu] = np.conj(np.swapaxes(gauge[mu, 1 - p, nidx], -1, -2))
    return out


def rel_err(a, b):
    """max over RHS of ||a_r - b_r|| / ||b_r|| (leading axis = RHS)."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return float(np.max(num / np.where(den == 0, 1, den)))


# ---------------------------------------------------------------- GPU plumbing
def dev_tensor(arr, dev):
    """numpy complex array -> torch device tensor of (re, im) pairs."""
    import torch

    a = np.ascontiguousarray(arr)
    if np.iscomplexobj(a):
        a = a.view(np.float64 if a.dtype == np.complex128 else np.float32).reshape(a.shape + (2,))
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def host_complex(t):
    """torch tensor of (re, im) pairs -> numpy complex."""
    import torch

    a = t.detach().to("cpu")
    if a.dtype == torch.float16:
        a = a.float()
    if a.dtype == torch.bfloat16:
        a = a.float()
    a = a.numpy()
    return a[..., 0] + 1j * a[..., 1]


def empty_prec(shape, prec, dev):
    """Device buffer of complex numbers in a library precision code."""
    import torch

    dt = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}[prec]
    return torch.empty(tuple(shape) + (2,), dtype=dt, device=dev)


def to_prec(arr, prec, dev):
    import torch

    t = dev_tensor(arr.astype(np.complex128), dev)
    dt = {0: torch.float16, 1: torch.float32, 2: torch.float64, 4: torch.bfloat16}[prec]
    return t.to(dt)


def sigma(mu, nu):
    """sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu]."""
    return 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])


def gauge_transform(gauge, L, om):
    """U_mu(x) -> om(x) U_mu(x) om(x+mu)^+ for om [2][S][N][N] (per parity, site)."""
    xh, y, z, t = fields.site_coords(L)
    out = np.empty_like(gauge)
    for p in range(2):
        X = 2 * xh + ((p + y + z + t) & 1)
        for mu in range(4):
            c = [X.copy(), y.copy(), z.copy(), t.copy()]
            c[mu] = (c[mu] + 1) % L[mu]
            nidx = cb_index(L, *c)
            out[mu, p] = om[p] @ gauge[mu, p] @ np.conj(np.swapaxes(om[1 - p, nidx], -1, -2))
    return out


def textbook_clover(gauge, L, coeff):
    """A(x) = 1 + coeff sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x) from full
    coordinates, F = (Q - Q^+)/(8i), Q the four plaquettes of the plane
    around x; returned as the full 4N x 4N matrix per site, [2][S][4N][4N]
    with index spin * N + colour."""
    N = gauge.shape[-1]
    S = fields.half_volume(L)
    xh, y, z, t = fields.site_coords(L)
    out = np.zeros((2, S, 4 * N, 4 * N), dtype=np.complex128)

    def dag(m):
        return np.conj(np.swapaxes(m, -1, -2))

    for p in range(2):
        X = 2 * xh + ((p + y + z + t) & 1)
        base = [X, y, z, t]

        def U(mu, shift):
            c = [(base[d] + shift[d]) % L[d] for d in range(4)]
            par = (c[0] + c[1] + c[2] + c[3]) & 1
            idx = cb_index(L, *c)
            return gauge[mu, par, idx]

        A = np.tile(np.eye(4 * N, dtype=np.complex128), (S, 1, 1))
        for mu in range(4):
            for nu in range(mu + 1, 4):
                e = lambda m, n: [m * (d == mu) + n * (d == nu) for d in range(4)]
                Q = U(mu, e(0, 0)) @ U(nu, e(1, 0)) @ dag(U(mu, e(0, 1))) @ dag(U(nu, e(0, 0)))
                Q = Q + U(nu, e(0, 0)) @ dag(U(mu, e(-1, 1))) @ dag(U(nu, e(-1, 0))) @ U(mu, e(-1, 0))
                Q = Q + dag(U(mu, e(-1, 0))) @ dag(U(nu, e(-1, -1))) @ U(mu, e(-1, -1)) @ U(nu, e(0, -1))
                Q = Q + dag(U(nu, e(0, -1))) @ U(mu, e(0, -1)) @ U(nu, e(1, -1)) @ dag(U(mu, e(0, 0)))
                F = (Q - dag(Q)) / 8j
                A = A + coeff * np.einsum("ab,sij->saibj", sigma(mu, nu), F).reshape(S, 4 * N, 4 * N)
        out[p] = A
    return out


def clover_blocks(A, N):
    """[2][S][4N][4N] -> the two diagonal spin-pair blocks [2][S][2][2N][2N]."""
    B = 2 * N
    return np.stack([A[..., :B, :B], A[..., B:, B:]], axis=2)
