"""CPU transcription of the HIP staggered kernel's Hines phase (test only).

Runs the chain walker of ``bc_stagger.hip`` (delta-V form, chain-major slots,
diagonal rebuilt from the zero row sums) and the transposed walker of
``bc_adjoint.hip`` over a packed ``StaggerLayout`` in NumPy, so that layout,
packing and both solves can be checked against the oracle without a GPU.
"""

import numpy as np


def _pack_ranges(L, p):
    P = L.pack_array
    return P[p, 2], P[p, 3], P[p, 4], P[p, 5], P[p, 6], P[p, 7]


def hines_delta(L, dt, bd_in, br_in, xv, dtype=np.float64):
    """bd_in/br_in/xv: per global slot.  Returns dV per global slot."""
    tb = L.dt_tables(dt)
    LO, UP = tb["lo"].astype(dtype), tb["up"].astype(dtype)
    bd = bd_in.astype(dtype).copy()
    br = br_in.astype(dtype).copy()
    xv = xv.astype(dtype)
    for p in range(L.n_packs):
        row0, nrows, ch0, nch, lv0, nlv = _pack_ranges(L, p)
        kd = np.zeros(nch, dtype)
        kr = np.zeros(nch, dtype)
        for lev in range(nlv):
            for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
                gk = ch0 + k
                j0, j1 = L.chain_off[gk] - row0, L.chain_off[gk + 1] - row0
                tp = L.chain_top[gk]
                e, r = bd[row0 + j0], br[row0 + j0]
                for q in range(L.kid_off[gk], L.kid_off[gk + 1]):
                    e -= kd[L.kids[q]]
                    r -= kr[L.kids[q]]
                vr = xv[row0 + j0]
                j = j0
                while True:
                    pr = j + 1 if j + 1 < j1 else tp
                    vp = xv[row0 + pr] if pr >= 0 else vr
                    lo, up = LO[row0 + j], UP[row0 + j]
                    dl = vr - vp
                    r = r + lo * dl
                    inv = dtype(1) / (e - lo)
                    bd[row0 + j] = inv
                    br[row0 + j] = r
                    kdc, krc = up * e * inv, up * r * inv
                    j += 1
                    if j >= j1:
                        break
                    vr = vp
                    e = bd[row0 + j] - kdc
                    r = (br[row0 + j] - krc) - up * dl
                kd[k] = kdc
                kr[k] = krc + up * dl
        _backsub(L, p, bd, br, LO)
    return br


def hines_transposed(L, dt, bd_in, b_in, dtype=np.float64):
    """Solves M^T y = b with M's diagonal bd_in (+ axial part) per slot."""
    tb = L.dt_tables(dt)
    LO, UP = tb["lo"].astype(dtype), tb["up"].astype(dtype)
    bd = bd_in.astype(dtype).copy()
    br = b_in.astype(dtype).copy()
    for p in range(L.n_packs):
        row0, nrows, ch0, nch, lv0, nlv = _pack_ranges(L, p)
        kd = np.zeros(nch, dtype)
        kr = np.zeros(nch, dtype)
        for lev in range(nlv):
            for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
                gk = ch0 + k
                j0, j1 = L.chain_off[gk] - row0, L.chain_off[gk + 1] - row0
                e, r = bd[row0 + j0], br[row0 + j0]
                for q in range(L.kid_off[gk], L.kid_off[gk + 1]):
                    e -= kd[L.kids[q]]
                    r -= kr[L.kids[q]]
                j = j0
                while True:
                    lo, up = LO[row0 + j], UP[row0 + j]
                    inv = dtype(1) / (e - lo)
                    bd[row0 + j] = inv
                    br[row0 + j] = r
                    kdc, krc = up * e * inv, lo * r * inv
                    j += 1
                    if j >= j1:
                        break
                    e = bd[row0 + j] - kdc
                    r = br[row0 + j] - krc
                kd[k] = kdc
                kr[k] = krc
        _backsub(L, p, bd, br, UP)
    return br


def hines_transposed_composite(L, dt, bd_in, b_in, dtype=np.float64):
    """The adjoint kernel's transposed elimination through chain composites:
    pass 1 composes every chain's rows into (a, b, h, c, f, g), rescaled to
    k = 1 after each group of four rows; pass 2 applies one composite per
    chain, level by level, to its kids' summed contributions (first kid,
    second kid, then the rest); pass 3 eliminates every chain's rows from
    that sum.  Solves M^T y = b like ``hines_transposed``."""
    tb = L.dt_tables(dt)
    LO, UP = tb["lo"].astype(dtype), tb["up"].astype(dtype)
    bd = bd_in.astype(dtype).copy()
    br = b_in.astype(dtype).copy()
    for p in range(L.n_packs):
        row0, nrows, ch0, nch, lv0, nlv = _pack_ranges(L, p)
        comp = np.zeros((nch, 6), dtype)
        kd = np.zeros(nch, dtype)
        kr = np.zeros(nch, dtype)

        def rows(k):
            gk = ch0 + k
            return L.chain_off[gk] - row0, L.chain_off[gk + 1] - row0

        def kid_sum(k):
            gk = ch0 + k
            sd = sr = dtype(0)
            for q in range(L.kid_off[gk], L.kid_off[gk + 1]):
                sd = sd + kd[L.kids[q]]
                sr = sr + kr[L.kids[q]]
            return sd, sr

        for k in range(nch):  # pass 1
            j0, j1 = rows(k)
            a, b, h, kk, c, f, g = (dtype(v) for v in (1, 0, 0, 1, 0, 1, 0))
            for j in range(j0, j1):
                d, r, lo, up = bd[row0 + j], br[row0 + j], LO[row0 + j], UP[row0 + j]
                a, b, c, g, h, kk, f = (up * (d * h - a), up * (d * kk - b), lo * (r * h - c), lo * (r * kk - g),
                                        (d - lo) * h - a, (d - lo) * kk - b, -lo * f)
                if (j - j0) % 4 == 3 or j == j1 - 1:
                    ik = dtype(1) / kk
                    a, b, h, c, f, g, kk = a * ik, b * ik, h * ik, c * ik, f * ik, g * ik, dtype(1)
            comp[k] = (a, b, h, c, f, g)
            if L.kid_off[ch0 + k + 1] == L.kid_off[ch0 + k]:
                kd[k], kr[k] = b, g
        for lev in range(1, nlv):  # pass 2
            for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
                a, b, h, c, f, g = comp[k]
                sd, sr = kid_sum(k)
                iz = dtype(1) / (h * sd + 1)
                kd[k], kr[k] = (a * sd + b) * iz, (c * sd + f * sr + g) * iz
        for k in range(nch):  # pass 3
            j0, j1 = rows(k)
            sd, sr = kid_sum(k)
            for j in range(j0, j1):
                e = bd[row0 + j] - sd
                r = br[row0 + j] - sr
                lo, up = LO[row0 + j], UP[row0 + j]
                inv = dtype(1) / (e - lo)
                bd[row0 + j] = inv
                br[row0 + j] = r
                sd, sr = up * e * inv, lo * r * inv
        _backsub(L, p, bd, br, UP)
    return br


def _backsub(L, p, bd, br, coef):
    row0, nrows, ch0, nch, lv0, nlv = _pack_ranges(L, p)
    for lev in range(nlv - 1, -1, -1):
        for k in range(L.level_off[lv0 + lev], L.level_off[lv0 + lev + 1]):
            gk = ch0 + k
            j0, j1 = L.chain_off[gk] - row0, L.chain_off[gk + 1] - row0
            tp = L.chain_top[gk]
            xp = br[row0 + tp] if tp >= 0 else 0.0
            for j in range(j1 - 1, j0 - 1, -1):
                x = (br[row0 + j] - coef[row0 + j] * xp) * bd[row0 + j]
                br[row0 + j] = x
                xp = x


def cv_rows(L):
    """Global slot of every CV (pack order)."""
    info = L.cv_info.astype(np.int64)
    rows = np.empty(L.n_cv_total, dtype=np.int64)
    P = L.pack_array
    for p in range(L.n_packs):
        cv0, ncv, row0 = P[p, 0], P[p, 1], P[p, 2]
        rows[cv0:cv0 + ncv] = row0 + (info[cv0:cv0 + ncv] & 0xFFFF)
    return rows


def bnd_rows(L):
    P = L.pack_array
    out = np.empty(len(L.bnd_info), dtype=np.int64)
    for p in range(L.n_packs):
        b0, nb, row0 = P[p, 8], P[p, 9], P[p, 2]
        out[b0:b0 + nb] = row0 + (L.bnd_info[b0:b0 + nb].astype(np.int64) & 0x7FFFFFFF)
    return out
