"""CPU model of the hypercube lane map (mrhs_qcu_amd/csrc/dslash_hcube.h):
for every lane of a launch, the output site, its own inner input, the three
corner partners (lane distance Pl, 2 Pl, 4 Pl), the outer neighbours, link
slots, hop signs and boundary-skip flags, restated from the kernel's index
arithmetic and checked against the canonical 8-direction neighbour table of
the site-major kernel (dslash_kernel.h site_geometry; reference
include/basic_data/qcu_point.cuh:33-88): every output site covered once per
RHS slot, every (direction, sign, neighbour, link slot, live) term present
exactly once, partners inside the same wavefront and block with the same
slot, every index in range, and the LDS link block of each lane inside its
workgroup's staged sites. The GPU parity of the kernels themselves is
tests/test_hcube_gpu.py."""
import itertools

import pytest


def decode(hb, A):
    subLx, bLy, zb, tb, nzb = A["subLx"], A["bLy"], A["zblock"], A["tblock"], A["nzb"]
    if A.get("col"):
        # column order: ((column * nT + T) * slab + (z_in * cy + y_in) * sx + x_in),
        # column = (z column * ny + y column) * nx + x column (x split in nx)
        cy, cz, nx = (tuple(A["col"]) + (1,))[:3]
        sx = subLx // nx
        q, b = divmod(hb, sx * cy * cz)
        colx, T = divmod(q, A["Lt"] // 2)
        col, cx_i = divmod(colx, nx)
        r, xin = divmod(b, sx)
        zin, yin = divmod(r, cy)
        cz_i, cy_i = divmod(col, bLy // cy)
        return cx_i * sx + xin, cy_i * cy + yin, cz_i * cz + zin, T
    if A["ystream"]:
        q, xh = divmod(hb, subLx)
        q2, zin = divmod(q, zb)
        q3, tin = divmod(q2, tb)
        q4, Y = divmod(q3, bLy)
        tout, zout = divmod(q4, nzb)
    else:
        q, r = divmod(hb, subLx * bLy)
        Y, xh = divmod(r, subLx)
        q2, zin = divmod(q, zb)
        q3, tin = divmod(q2, tb)
        tout, zout = divmod(q3, nzb)
    return xh, Y, zout * zb + zin, tout * tb + tin


def canonical(idx, A, p):
    subLx, Ly, Lz, Lt = A["subLx"], A["Ly"], A["Lz"], A["Lt"]
    xh = idx % subLx
    y = idx // subLx % Ly
    z = idx // (subLx * Ly) % Lz
    t = idx // (subLx * Ly * Lz)
    eo = (y + z + t) & 1
    xf = (xh + 1) % subLx if eo != p else xh
    xb = (xh - 1) % subLx if eo == p else xh
    sy, sz, st = subLx, subLx * Ly, subLx * Ly * Lz
    nb = [idx - xh + xf, idx - xh + xb,
          idx + (-(Ly - 1) * sy if y + 1 == Ly else sy), idx + ((Ly - 1) * sy if y == 0 else -sy),
          idx + (-(Lz - 1) * sz if z + 1 == Lz else sz), idx + ((Lz - 1) * sz if z == 0 else -sz),
          idx + (-(Lt - 1) * st if t + 1 == Lt else st), idx + ((Lt - 1) * st if t == 0 else -st)]
    sk = A["skip"]
    live = [not (sk[0] and xh == subLx - 1 and p != eo), not (sk[0] and xh == 0 and p == eo),
            not (sk[1] and y == Ly - 1), not (sk[1] and y == 0),
            not (sk[2] and z == Lz - 1), not (sk[2] and z == 0),
            not (sk[3] and t == Lt - 1), not (sk[3] and t == 0)]
    return nb, live


def col_remap(hw, A, B, P):
    """Logical workgroup of hardware workgroup hw in column order (None: idle)."""
    cy, cz, nx = (tuple(A["col"]) + (1,))[:3]
    ncol = (A["bLy"] // cy) * (A["bLz"] // cz) * nx
    grp = (A["Lt"] // 2) * (A["subLx"] // nx * cy * cz // (B // (8 * P)))
    i, x = hw >> 3, hw & 7
    g = i // grp
    col = 8 * g + x
    return col * grp + (i - g * grp) if col < ncol else None


def check_launch(L, P, B, tb, zb, ystream, p, skip, col=None):
    subLx, Ly, Lz, Lt = L[0] // 2, L[1], L[2], L[3]
    S = subLx * Ly * Lz * Lt
    A = dict(subLx=subLx, Ly=Ly, Lz=Lz, Lt=Lt, bLy=Ly // 2, bLz=Lz // 2, zblock=zb or Lz // 2, tblock=tb,
             nzb=(Lz // 2) // (zb or Lz // 2), ystream=ystream, skip=skip, col=col)
    lp = P.bit_length() - 1
    lpl, gb = min(lp, 3), 3 + lp
    nlanes = S * P
    lanes = []
    for gid in range(nlanes):
        rr = gid & ((1 << gb) - 1)
        hi, r2 = rr >> (3 + lpl), rr & ((1 << (3 + lpl)) - 1)
        cn = r2 >> lpl
        slot = (hi << lpl) | (r2 & ((1 << lpl) - 1))
        sw = ((gid >> gb) << 3) | cn
        xh, Y, Z, T = decode(sw >> 3, A)
        y, z, t = 2 * Y + (cn & 1), 2 * Z + ((cn >> 1) & 1), 2 * T + (cn >> 2)
        lanes.append((sw, slot, cn, xh, y, z, t, ((t * Lz + z) * Ly + y) * subLx + xh))
    seen = set()
    for (sw, slot, cn, xh, y, z, t, idx) in lanes:
        assert 0 <= idx < S and slot < P
        assert (idx, slot) not in seen
        seen.add((idx, slot))
    assert len(seen) == S * P
    for blk in range((nlanes + B - 1) // B):
        lane0, end = blk * B, min(blk * B + B, nlanes)
        nsite = (end - lane0) >> lp
        assert nsite * P == end - lane0 and nsite <= (B - 1) // P + 2
        for gid in range(lane0, end):
            assert 0 <= lanes[gid][0] - (lane0 >> lp) < nsite
    Pl = 1 << lpl
    for gid, (sw, slot, cn, xh, y, z, t, idx) in enumerate(lanes):
        e = p ^ ((y + z + t) & 1)
        sg = [-1 if e else 1, -1 if cn & 1 else 1, -1 if cn & 2 else 1, -1 if cn & 4 else 1]
        dy, dz = subLx, subLx * Ly
        dt = dz * Lz
        outer = [(idx - xh if xh + 1 == subLx else idx + 1) if e else (idx + subLx - 1 if xh == 0 else idx - 1),
                 (idx - (Ly - 1) * dy if y + 1 == Ly else idx + dy) if cn & 1 else (idx + (Ly - 1) * dy if y == 0 else idx - dy),
                 (idx - (Lz - 1) * dz if z + 1 == Lz else idx + dz) if cn & 2 else (idx + (Lz - 1) * dz if z == 0 else idx - dz),
                 (idx - (Lt - 1) * dt if t + 1 == Lt else idx + dt) if cn & 4 else (idx + (Lt - 1) * dt if t == 0 else idx - dt)]
        live = [not (skip[0] and (xh == subLx - 1 if e else xh == 0)),
                not (skip[1] and (y == Ly - 1 if cn & 1 else y == 0)),
                not (skip[2] and (z == Lz - 1 if cn & 2 else z == 0)),
                not (skip[3] and (t == Lt - 1 if cn & 4 else t == 0))]
        terms = {(0, sg[0], idx, 4 if e else 0, True)}
        for mu in (1, 2, 3):
            partner = gid ^ (Pl << (mu - 1))
            psw, pslot, pcn = lanes[partner][:3]
            assert partner >> 6 == gid >> 6 and psw >> 3 == sw >> 3 and pslot == slot
            # the partner projects with minus its own sign; this lane applies its own
            assert -(-1 if pcn & (1 << (mu - 1)) else 1) == sg[mu]
            terms.add((mu, sg[mu], lanes[partner][7], (4 + mu) if cn & (1 << (mu - 1)) else mu, True))
        for mu in range(4):
            fwd_out = -sg[mu] > 0
            terms.add((mu, -sg[mu], outer[mu], mu if fwd_out else 4 + mu, live[mu]))
        nb, lv = canonical(idx, A, p)
        want = {(D >> 1, 1 if D % 2 == 0 else -1, nb[D], (D >> 1) if D % 2 == 0 else 4 + (D >> 1), lv[D])
                for D in range(8)}
        assert terms == want, (L, P, B, tb, zb, ystream, p, gid)
    if col:
        # The padded hardware grid maps onto every logical workgroup once; XCD
        # hw % 8 works on columns congruent to it, one slab after another with T
        # ascending, and the blocks of a slab stay inside its column.
        cy, cz, nx = (tuple(col) + (1,))[:3]
        ncol = (Ly // 2 // cy) * (Lz // 2 // cz) * nx
        nb = nlanes // B
        grp = (Lt // 2) * (subLx // nx * cy * cz // (B // (8 * P)))
        assert nlanes % B == 0 and nb == ncol * grp
        ngrid = (ncol + 7) // 8 * 8 * grp
        got = [col_remap(hw, A, B, P) for hw in range(ngrid)]
        live = [g for g in got if g is not None]
        assert sorted(live) == list(range(nb))
        slab_wg = grp // (Lt // 2)
        for x in range(8):
            seq = [g for hw, g in enumerate(got) if hw % 8 == x and g is not None]
            cols = [g // grp for g in seq]
            assert all(c % 8 == x for c in cols)
            for a_, b_ in zip(seq, seq[1:]):
                if a_ // grp == b_ // grp:
                    assert (b_ % grp) // slab_wg >= (a_ % grp) // slab_wg
        for blk in range(nb):
            c, cxi = divmod(blk // grp, nx)
            cyi, czi = c % (Ly // 2 // cy), c // (Ly // 2 // cy)
            sx = subLx // nx
            for gid in range(blk * B, blk * B + B):
                X, Y, Z = lanes[gid][3], lanes[gid][4] // 2, lanes[gid][5] // 2
                assert cyi * cy <= Y < (cyi + 1) * cy and czi * cz <= Z < (czi + 1) * cz
                assert cxi * sx <= X < (cxi + 1) * sx


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (2, 2, 2, 2), (8, 4, 6, 2), (6, 2, 4, 8)])
def test_hypercube_lane_map(L):
    n = 0
    for P, B in ((1, 64), (2, 128), (4, 64), (8, 128), (8, 256), (16, 128), (32, 256)):
        bt, bz = L[3] // 2, L[2] // 2
        for tb, zb in itertools.product([d for d in (1, 2, 3, 4) if bt % d == 0], [0] + [d for d in (1, 3) if bz % d == 0]):
            for ystream, p, skip in ((0, 0, (0, 0, 0, 0)), (1, 1, (1, 1, 1, 1)), (1, 0, (1, 0, 1, 0))):
                check_launch(L, P, B, tb, zb, ystream, p, skip)
                n += 1
    assert n > 0


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (8, 4, 6, 2), (6, 8, 4, 4), (4, 12, 12, 4)])
def test_column_order_lane_map(L):
    """Column order (traversal 6): the same lane-level checks, plus the
    hardware-to-logical workgroup map of the padded grid."""
    n = 0
    for P, B in ((1, 64), (2, 128), (8, 128), (8, 256), (16, 128)):
        by, bz = L[1] // 2, L[2] // 2
        for cy, cz in itertools.product([d for d in (1, 2, 3) if by % d == 0], [d for d in (1, 2, 3) if bz % d == 0]):
            for nx in (1, 2):
                if (L[0] // 2) % nx or (L[0] // 2 // nx * cy * cz) % (B // (8 * P)):
                    continue
                for p, skip in ((0, (0, 0, 0, 0)), (1, (1, 0, 1, 1))):
                    check_launch(L, P, B, 1, 0, 0, p, skip, col=(cy, cz, nx))
                    n += 1
    assert n > 0


def test_site_major_column_order_map():
    """Column order in sites (traversal 4, dslash_kernel.h decode_site and
    col_remap): logical site ((column * Lt + t) * slab + (z_in * cy + y_in) *
    subLx + xh); the padded grid of workgroups maps onto every logical
    workgroup once, each XCD on columns congruent to it with t ascending, and
    the sites of a workgroup inside its column."""
    n = 0
    for L, P, B in (((8, 4, 6, 8), 2, 128), ((4, 8, 12, 6), 8, 128), ((6, 6, 4, 4), 1, 64), ((4, 12, 12, 8), 4, 256)):
        subLx, Ly, Lz, Lt = L[0] // 2, L[1], L[2], L[3]
        S = subLx * Ly * Lz * Lt
        for cy in [d for d in (1, 2, 3, 4, 6) if Ly % d == 0]:
            for cz in [d for d in (1, 2, 3, 4, 6) if Lz % d == 0]:
                col_lanes = subLx * cy * cz * Lt * P
                if col_lanes % B:
                    continue
                ncy, ncol = Ly // cy, (Ly // cy) * (Lz // cz)
                grp = col_lanes // B
                nb = S * P // B
                assert nb == ncol * grp
                seen = set()
                for sl in range(S):
                    q, b = divmod(sl, subLx * cy * cz)
                    col, t = divmod(q, Lt)
                    r, xh = divmod(b, subLx)
                    zin, yin = divmod(r, cy)
                    czi, cyi = divmod(col, ncy)
                    y, z = cyi * cy + yin, czi * cz + zin
                    assert 0 <= y < Ly and 0 <= z < Lz and t < Lt
                    seen.add(((t * Lz + z) * Ly + y) * subLx + xh)
                assert len(seen) == S
                ngrid = (ncol + 7) // 8 * 8 * grp
                got = []
                for hw in range(ngrid):
                    i, x = hw >> 3, hw & 7
                    g = i // grp
                    c = 8 * g + x
                    got.append(c * grp + (i - g * grp) if c < ncol else None)
                assert sorted(v for v in got if v is not None) == list(range(nb))
                for hw, v in enumerate(got):
                    if v is not None:
                        assert (v // grp) % 8 == hw % 8
                        # every lane of the workgroup is a site of column v // grp
                        for gid in (v * B, v * B + B - 1):
                            assert (gid // P) // (subLx * cy * cz * Lt) == v // grp
                n += 1
    assert n > 10
